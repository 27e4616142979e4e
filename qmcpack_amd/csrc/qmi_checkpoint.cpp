// Checkpoint container: writer, validating reader and the RNG state form (host only).
#include "qmi_checkpoint.hpp"

#include <cstring>
#include <iterator>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace qmi
{
namespace
{
constexpr char MAGIC[8]       = {'Q', 'M', 'I', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t VERSION    = 1;
constexpr size_t HEADER_BYTES = 16;

size_t type_size(uint32_t t)
{
  switch (t)
  {
  case CKPT_F64: return 8;
  case CKPT_I64: return 8;
  case CKPT_U32: return 4;
  default: return 0;
  }
}

template<class T>
void put(std::vector<char>& b, const T& v)
{
  const char* p = reinterpret_cast<const char*>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}
} // namespace

CkptWriter::CkptWriter()
{
  bytes.insert(bytes.end(), MAGIC, MAGIC + 8);
  put(bytes, VERSION);
  put(bytes, uint32_t(0));
}

void CkptWriter::add(const std::string& name, CkptType type, const std::vector<uint64_t>& dims, const void* data)
{
  put(bytes, uint32_t(name.size()));
  bytes.insert(bytes.end(), name.begin(), name.end());
  put(bytes, uint32_t(type));
  put(bytes, uint32_t(dims.size()));
  uint64_t n = 1;
  for (uint64_t d : dims)
  {
    put(bytes, d);
    n *= d;
  }
  const char* p = static_cast<const char*>(data);
  if (n)
    bytes.insert(bytes.end(), p, p + n * type_size(type));
  count++;
}

void CkptWriter::add_f64(const std::string& name, const std::vector<double>& v)
{
  add(name, CKPT_F64, {uint64_t(v.size())}, v.data());
}

void CkptWriter::add_i64(const std::string& name, const std::vector<int64_t>& v)
{
  add(name, CKPT_I64, {uint64_t(v.size())}, v.data());
}

const std::vector<char>& CkptWriter::finish()
{
  std::memcpy(bytes.data() + 12, &count, sizeof(count));
  return bytes;
}

size_t CkptDataset::size() const
{
  size_t n = 1;
  for (uint64_t d : dims)
    n *= d;
  return n;
}

CkptReader::CkptReader(const void* buf, size_t n)
{
  const char* b = static_cast<const char*>(buf);
  if (!b || n < HEADER_BYTES || std::memcmp(b, MAGIC, 8) != 0)
    throw std::runtime_error("not a walker checkpoint (bad magic)");
  uint32_t version, count;
  std::memcpy(&version, b + 8, 4);
  std::memcpy(&count, b + 12, 4);
  if (version != VERSION)
    throw std::runtime_error("unsupported checkpoint format version " + std::to_string(version));
  size_t off = HEADER_BYTES;
  auto take  = [&](void* dst, size_t k) {
    if (k > n - off)
      throw std::runtime_error("truncated checkpoint");
    std::memcpy(dst, b + off, k);
    off += k;
  };
  for (uint32_t i = 0; i < count; i++)
  {
    uint32_t len, type, rank;
    take(&len, 4);
    if (len > n - off || len > 4096)
      throw std::runtime_error("truncated checkpoint");
    std::string name(b + off, len);
    off += len;
    take(&type, 4);
    take(&rank, 4);
    if (!type_size(type) || rank > 8)
      throw std::runtime_error("malformed checkpoint dataset '" + name + "'");
    CkptDataset d{CkptType(type), std::vector<uint64_t>(rank), nullptr};
    uint64_t elems = 1;
    for (uint32_t r = 0; r < rank; r++)
    {
      take(&d.dims[r], 8);
      if (d.dims[r] && elems > std::numeric_limits<uint64_t>::max() / d.dims[r])
        throw std::runtime_error("malformed checkpoint dataset '" + name + "'");
      elems *= d.dims[r];
    }
    const size_t es = type_size(type);
    if (elems > (n - off) / es)
      throw std::runtime_error("truncated checkpoint");
    d.data = b + off;
    off += elems * es;
    sets.emplace_back(std::move(name), std::move(d));
  }
  if (off != n)
    throw std::runtime_error("trailing bytes after the checkpoint's datasets");
}

const CkptDataset* CkptReader::find(const std::string& name) const
{
  for (auto& s : sets)
    if (s.first == name)
      return &s.second;
  return nullptr;
}

const CkptDataset& CkptReader::need(const std::string& name, CkptType type, long long count) const
{
  const CkptDataset* d = find(name);
  if (!d)
    throw std::runtime_error("checkpoint has no dataset '" + name + "'");
  if (d->type != type || (count >= 0 && d->size() != size_t(count)))
    throw std::runtime_error("checkpoint dataset '" + name + "' has the wrong type or size");
  return *d;
}

std::vector<double> CkptReader::f64(const std::string& name, long long count) const
{
  const CkptDataset& d = need(name, CKPT_F64, count);
  std::vector<double> v(d.size());
  if (!v.empty())
    std::memcpy(v.data(), d.data, v.size() * 8);
  return v;
}

std::vector<int64_t> CkptReader::i64(const std::string& name, long long count) const
{
  const CkptDataset& d = need(name, CKPT_I64, count);
  std::vector<int64_t> v(d.size());
  if (!v.empty())
    std::memcpy(v.data(), d.data, v.size() * 8);
  return v;
}

std::vector<uint32_t> mt_save(const std::mt19937& e)
{
  std::stringstream s;
  s << e;
  return std::vector<uint32_t>(std::istream_iterator<uint32_t>(s), std::istream_iterator<uint32_t>());
}

void mt_load(std::mt19937& e, const uint32_t* words, size_t n)
{
  if (n != mt_state_size())
    throw std::runtime_error("random stream state of " + std::to_string(n) + " numbers, expected " +
                             std::to_string(mt_state_size()));
  // a position beyond the 624 words would make the engine read past its state
  if (n == 625 && words[624] > 624)
    throw std::runtime_error("invalid random stream state (position " + std::to_string(words[624]) + ")");
  std::stringstream s;
  for (size_t i = 0; i < n; i++)
    s << words[i] << ' ';
  std::mt19937 t;
  if (!(s >> t))
    throw std::runtime_error("invalid random stream state");
  e = t;
}

size_t mt_state_size()
{
  static const size_t n = mt_save(std::mt19937()).size();
  return n;
}
} // namespace qmi
