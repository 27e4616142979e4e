// Walker / random-number / branch checkpoint container (host only).
//
// The reference writes a restart file with HDF5 (HDFWalkerOutput::dump, Particle/HDFWalkerOutput.cpp:
// 81-108: state_0/{block, number_of_walkers, walker_partition, walkers [nw][np][3],
// walker_weights [nw]}; RandomNumberControl::write_rank_0, Utilities/RandomNumberControl.cpp:382-:
// nprocs_nthreads_statesize, random/<engine> [streams][state_size]; BranchIO::write,
// QMCDrivers/BranchIO.cpp:109-160: branch parameters and histories).  The same named datasets are
// held here in a flat little-endian container, so that the library needs no HDF5:
//
//   char magic[8] = "QMICKPT1"; u32 format version (1); u32 dataset count; then per dataset
//   u32 name length, name bytes, u32 type (1 f64, 2 i64, 3 u32), u32 rank, u64 dims[rank],
//   the row-major payload.
//
// A random stream's state is the text form of std::mt19937 read back as numbers, as
// StdRandom::save / load do (Utilities/StdRandom.cpp): 624 state words and, with libstdc++, the
// position in them (625 numbers).
#pragma once
#include <cstddef>
#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace qmi
{
enum CkptType : uint32_t
{
  CKPT_F64 = 1,
  CKPT_I64 = 2,
  CKPT_U32 = 3
};

struct CkptWriter
{
  std::vector<char> bytes;
  uint32_t count = 0;
  CkptWriter();
  void add(const std::string& name, CkptType type, const std::vector<uint64_t>& dims, const void* data);
  void add_f64(const std::string& name, const std::vector<double>& v);
  void add_i64(const std::string& name, const std::vector<int64_t>& v);
  // closes the container (writes the dataset count)
  const std::vector<char>& finish();
};

struct CkptDataset
{
  CkptType type;
  std::vector<uint64_t> dims;
  const char* data;
  size_t size() const; // elements
};

// Parses and validates a container; throws std::runtime_error with a message on malformed input
struct CkptReader
{
  std::vector<std::pair<std::string, CkptDataset>> sets;
  CkptReader(const void* buf, size_t n);
  const CkptDataset* find(const std::string& name) const;
  // the named dataset with the given type and element count (count < 0: any), else throws
  const CkptDataset& need(const std::string& name, CkptType type, long long count = -1) const;
  std::vector<double> f64(const std::string& name, long long count = -1) const;
  std::vector<int64_t> i64(const std::string& name, long long count = -1) const;
};

// StdRandom::save / load: the engine state as numbers
std::vector<uint32_t> mt_save(const std::mt19937& e);
void mt_load(std::mt19937& e, const uint32_t* words, size_t n);
size_t mt_state_size();
} // namespace qmi
