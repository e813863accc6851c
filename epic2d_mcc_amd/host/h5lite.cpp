// Minimal HDF5 writer; the structures follow the HDF5 File Format
// Specification (version 3.0): superblock version 2 (section II.A),
// version 2 object headers (IV.A.1.b), and the header messages Dataspace
// (0x0001, version 2), Link Info (0x0002), Datatype (0x0003, version 1),
// Fill Value (0x0005, version 3), Link (0x0006), Data Layout (0x0008,
// version 3, contiguous), Group Info (0x000A) and Attribute (0x000C,
// version 3).  All multi-byte fields are little-endian; offsets and lengths
// are 8 bytes.
#include "h5lite.hpp"

#include <cstring>

namespace epic2d {

namespace {

constexpr uint64_t UNDEF = ~0ull;  // the format's undefined address
constexpr uint64_t SUPERBLOCK_BYTES = 48;

inline uint32_t rot(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

struct Buf {
  std::vector<uint8_t> b;
  void u8(unsigned v) { b.push_back((uint8_t)v); }
  void u16(unsigned v) {
    u8(v & 0xff);
    u8((v >> 8) & 0xff);
  }
  void u32(uint32_t v) {
    for (int k = 0; k < 4; k++) u8((v >> (8 * k)) & 0xff);
  }
  void u64(uint64_t v) {
    for (int k = 0; k < 8; k++) u8((unsigned)((v >> (8 * k)) & 0xff));
  }
  void bytes(const void* p, size_t n) {
    const uint8_t* q = (const uint8_t*)p;
    b.insert(b.end(), q, q + n);
  }
  void cat(const std::vector<uint8_t>& v) { b.insert(b.end(), v.begin(), v.end()); }
};

// Datatype messages (class and version byte, three class bit-field bytes,
// size, properties).
std::vector<uint8_t> type_f64() {
  Buf t;
  t.u8(0x11);  // version 1, class 1: floating point
  t.u8(0x20);  // little-endian, no padding, mantissa normalisation 2 (implied leading bit)
  t.u8(63);    // sign bit position
  t.u8(0);
  t.u32(8);
  t.u16(0);   // bit offset
  t.u16(64);  // bit precision
  t.u8(52);   // exponent location
  t.u8(11);   // exponent size
  t.u8(0);    // mantissa location
  t.u8(52);   // mantissa size
  t.u32(1023);
  return t.b;
}

std::vector<uint8_t> type_i32() {
  Buf t;
  t.u8(0x10);  // version 1, class 0: fixed point
  t.u8(0x08);  // little-endian, no padding, signed
  t.u8(0);
  t.u8(0);
  t.u32(4);
  t.u16(0);   // bit offset
  t.u16(32);  // bit precision
  return t.b;
}

std::vector<uint8_t> type_str(uint32_t size) {
  Buf t;
  t.u8(0x13);  // version 1, class 3: string
  t.u8(0x00);  // NUL-terminated, ASCII
  t.u8(0);
  t.u8(0);
  t.u32(size);
  return t.b;
}

std::vector<uint8_t> space_scalar() {
  Buf s;
  s.u8(2);  // version 2
  s.u8(0);  // rank 0
  s.u8(0);  // no maximum dimensions
  s.u8(0);  // scalar
  return s.b;
}

std::vector<uint8_t> space_simple(const std::vector<uint64_t>& dims) {
  Buf s;
  s.u8(2);
  s.u8((unsigned)dims.size());
  s.u8(0);
  s.u8(1);  // simple
  for (uint64_t d : dims) s.u64(d);
  return s.b;
}

// One header message of a version 2 object header: type, size, flags, data.
void message(Buf& h, unsigned type, const std::vector<uint8_t>& data, unsigned flags = 0) {
  h.u8(type);
  h.u16((unsigned)data.size());
  h.u8(flags);
  h.cat(data);
}

}  // namespace

uint32_t lookup3(const void* data, size_t length, uint32_t initval) {
  const uint8_t* k = (const uint8_t*)data;
  uint32_t a, b, c;
  a = b = c = 0xdeadbeefu + (uint32_t)length + initval;
  while (length > 12) {
    a += k[0] + ((uint32_t)k[1] << 8) + ((uint32_t)k[2] << 16) + ((uint32_t)k[3] << 24);
    b += k[4] + ((uint32_t)k[5] << 8) + ((uint32_t)k[6] << 16) + ((uint32_t)k[7] << 24);
    c += k[8] + ((uint32_t)k[9] << 8) + ((uint32_t)k[10] << 16) + ((uint32_t)k[11] << 24);
    // mix
    a -= c; a ^= rot(c, 4);  c += b;
    b -= a; b ^= rot(a, 6);  a += c;
    c -= b; c ^= rot(b, 8);  b += a;
    a -= c; a ^= rot(c, 16); c += b;
    b -= a; b ^= rot(a, 19); a += c;
    c -= b; c ^= rot(b, 4);  b += a;
    length -= 12;
    k += 12;
  }
  if (length == 0) return c;
  // the last 1..12 bytes, zero-extended
  uint8_t tail[12] = {0};
  std::memcpy(tail, k, length);
  a += tail[0] + ((uint32_t)tail[1] << 8) + ((uint32_t)tail[2] << 16) + ((uint32_t)tail[3] << 24);
  b += tail[4] + ((uint32_t)tail[5] << 8) + ((uint32_t)tail[6] << 16) + ((uint32_t)tail[7] << 24);
  c += tail[8] + ((uint32_t)tail[9] << 8) + ((uint32_t)tail[10] << 16) + ((uint32_t)tail[11] << 24);
  // final
  c ^= b; c -= rot(b, 14);
  a ^= c; a -= rot(c, 11);
  b ^= a; b -= rot(a, 25);
  c ^= b; c -= rot(b, 16);
  a ^= c; a -= rot(c, 4);
  b ^= a; b -= rot(a, 14);
  c ^= b; c -= rot(b, 24);
  return c;
}

H5Lite::H5Lite(const std::string& path) {
  f_ = std::fopen(path.c_str(), "wb");
  if (!f_) return;
  // the superblock's place, written last
  uint8_t zero[SUPERBLOCK_BYTES] = {0};
  if (std::fwrite(zero, 1, sizeof(zero), f_) != sizeof(zero)) failed_ = true;
  eof_ = SUPERBLOCK_BYTES;
  objs_.push_back(Obj{});  // the root group
}

H5Lite::~H5Lite() {
  if (f_) close();
}

int H5Lite::find(int parent, const std::string& name) const {
  if (parent < 0 || parent >= (int)objs_.size()) return -1;
  for (int k : objs_[parent].kids)
    if (objs_[k].name == name) return k;
  return -1;
}

int H5Lite::group(int parent, const std::string& name) {
  if (parent < 0 || parent >= (int)objs_.size() || !objs_[parent].is_group || name.empty() ||
      name.find('/') != std::string::npos)
    return -1;
  const int k = find(parent, name);
  if (k >= 0) return objs_[k].is_group ? k : -1;
  Obj g;
  g.name = name;
  g.parent = parent;
  objs_.push_back(g);
  objs_[parent].kids.push_back((int)objs_.size() - 1);
  return (int)objs_.size() - 1;
}

int H5Lite::dataset_f64(int parent, const std::string& name, const std::vector<uint64_t>& dims, const double* data) {
  if (!ok() || parent < 0 || parent >= (int)objs_.size() || !objs_[parent].is_group || name.empty() ||
      name.find('/') != std::string::npos || find(parent, name) >= 0 || dims.empty() || dims.size() > 4)
    return -1;
  uint64_t n = 1;
  for (uint64_t d : dims) n *= d;
  Obj d;
  d.is_group = false;
  d.name = name;
  d.parent = parent;
  d.dims = dims;
  d.data_size = n * sizeof(double);
  d.data_addr = n ? eof_ : UNDEF;
  if (n) {
    if (std::fwrite(data, sizeof(double), n, f_) != n) {
      failed_ = true;
      return -1;
    }
    eof_ += d.data_size;
  }
  objs_.push_back(d);
  objs_[parent].kids.push_back((int)objs_.size() - 1);
  return (int)objs_.size() - 1;
}

bool H5Lite::add_attr(int obj, Attr a) {
  if (obj < 0 || obj >= (int)objs_.size() || a.name.empty()) return false;
  for (const Attr& x : objs_[obj].attrs)
    if (x.name == a.name) return false;
  objs_[obj].attrs.push_back(std::move(a));
  return true;
}

bool H5Lite::attr_i32(int obj, const std::string& name, int32_t v) {
  Attr a{name, type_i32(), space_scalar(), {}};
  Buf d;
  d.u32((uint32_t)v);
  a.data = d.b;
  return add_attr(obj, std::move(a));
}

bool H5Lite::attr_f64(int obj, const std::string& name, double v) {
  Attr a{name, type_f64(), space_scalar(), {}};
  uint64_t bits;
  std::memcpy(&bits, &v, 8);
  Buf d;
  d.u64(bits);
  a.data = d.b;
  return add_attr(obj, std::move(a));
}

bool H5Lite::attr_str(int obj, const std::string& name, const std::string& s, uint32_t size) {
  if (size == 0) return false;
  Attr a{name, type_str(size), space_scalar(), std::vector<uint8_t>(size, 0)};
  std::memcpy(a.data.data(), s.data(), std::min<size_t>(size, s.size()));
  return add_attr(obj, std::move(a));
}

std::vector<uint8_t> H5Lite::header(const Obj& o, const std::vector<uint64_t>& addr) const {
  Buf m;  // the messages of chunk 0
  if (o.is_group) {
    Buf li;  // Link Info: no creation order, compact storage (no fractal heap, no name index)
    li.u8(0);
    li.u8(0);
    li.u64(UNDEF);
    li.u64(UNDEF);
    message(m, 0x0002, li.b);
    Buf gi;  // Group Info: defaults
    gi.u8(0);
    gi.u8(0);
    message(m, 0x000A, gi.b);
    for (int k : o.kids) {
      const std::string& nm = objs_[k].name;
      Buf ln;  // Link: hard link, ASCII name
      ln.u8(1);
      if (nm.size() < 256) {
        ln.u8(0x00);  // 1-byte name length
        ln.u8((unsigned)nm.size());
      } else {
        ln.u8(0x01);  // 2-byte name length
        ln.u16((unsigned)nm.size());
      }
      ln.bytes(nm.data(), nm.size());
      ln.u64(addr[k]);
      message(m, 0x0006, ln.b);
    }
  } else {
    message(m, 0x0001, space_simple(o.dims));
    message(m, 0x0003, type_f64(), 0x01);  // constant
    Buf fv;  // Fill Value: allocation late, fill written only if set, none defined
    fv.u8(3);
    fv.u8(0x0A);
    message(m, 0x0005, fv.b, 0x01);
    Buf ly;  // Data Layout: contiguous
    ly.u8(3);
    ly.u8(1);
    ly.u64(o.data_addr);
    ly.u64(o.data_size);
    message(m, 0x0008, ly.b);
  }
  for (const Attr& a : o.attrs) {
    Buf at;
    at.u8(3);
    at.u8(0);
    at.u16((unsigned)a.name.size() + 1);
    at.u16((unsigned)a.type.size());
    at.u16((unsigned)a.space.size());
    at.u8(0);  // ASCII name
    at.bytes(a.name.c_str(), a.name.size() + 1);
    at.cat(a.type);
    at.cat(a.space);
    at.cat(a.data);
    message(m, 0x000C, at.b);
  }
  Buf h;
  h.bytes("OHDR", 4);
  h.u8(2);     // version
  h.u8(0x12);  // chunk-0 size in 4 bytes; attribute phase-change values stored
  h.u16(1024);  // maximum compact attributes
  h.u16(1000);  // minimum dense attributes
  h.u32((uint32_t)m.b.size());
  h.cat(m.b);
  h.u32(lookup3(h.b.data(), h.b.size(), 0));
  return h.b;
}

bool H5Lite::close() {
  if (!f_) return false;
  bool good = !failed_;
  if (good) {
    // header sizes do not depend on the addresses: size, place, then encode
    std::vector<uint64_t> addr(objs_.size(), 0);
    std::vector<size_t> size(objs_.size());
    for (size_t k = 0; k < objs_.size(); k++) size[k] = header(objs_[k], addr).size();
    uint64_t at = eof_;
    for (size_t k = 0; k < objs_.size(); k++) {
      addr[k] = at;
      at += size[k];
    }
    for (size_t k = 0; k < objs_.size() && good; k++) {
      const std::vector<uint8_t> h = header(objs_[k], addr);
      good = std::fwrite(h.data(), 1, h.size(), f_) == h.size();
    }
    Buf sb;
    sb.bytes("\x89HDF\r\n\x1a\n", 8);
    sb.u8(2);  // superblock version
    sb.u8(8);  // size of offsets
    sb.u8(8);  // size of lengths
    sb.u8(0);  // file consistency flags
    sb.u64(0);      // base address
    sb.u64(UNDEF);  // no superblock extension
    sb.u64(at);     // end-of-file address
    sb.u64(addr[0]);  // root group object header
    sb.u32(lookup3(sb.b.data(), sb.b.size(), 0));
    good = good && std::fseek(f_, 0, SEEK_SET) == 0 && std::fwrite(sb.b.data(), 1, sb.b.size(), f_) == sb.b.size();
  }
  good = (std::fclose(f_) == 0) && good;
  f_ = nullptr;
  return good;
}

}  // namespace epic2d
