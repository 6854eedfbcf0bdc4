// CGNS (HDF5 flavour) writer for the driver's solution files; the tree it
// writes is the one zone_FD/src/src/cgnsio.cpp:3-101 produces through libcgns.
#include "cgns_writer.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace cfdhost {

namespace {

uint32_t le32(const unsigned char* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

}  // namespace

unsigned lookup3(const unsigned char* k, size_t length, unsigned initval) {
  auto rot = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
  uint32_t a, b, c;
  a = b = c = 0xdeadbeefu + (uint32_t)length + initval;
  while (length > 12) {
    a += le32(k);
    b += le32(k + 4);
    c += le32(k + 8);
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
  unsigned char tail[12] = {0};
  std::memcpy(tail, k, length);
  a += le32(tail);
  b += le32(tail + 4);
  c += le32(tail + 8);
  c ^= b; c -= rot(b, 14);
  a ^= c; a -= rot(c, 11);
  b ^= a; b -= rot(a, 25);
  c ^= b; c -= rot(b, 16);
  a ^= c; a -= rot(c, 4);
  b ^= a; b -= rot(a, 14);
  c ^= b; c -= rot(b, 24);
  return c;
}

namespace {

constexpr uint64_t UNDEF = ~0ull;
using Bytes = std::vector<unsigned char>;

void put(Bytes& b, uint64_t v, int n) {
  for (int i = 0; i < n; i++) b.push_back((unsigned char)(v >> (8 * i)));
}
void put(Bytes& b, const void* p, size_t n) {
  auto c = static_cast<const unsigned char*>(p);
  b.insert(b.end(), c, c + n);
}

// HDF5 datatype messages (class | version 1 << 4, class bits, size, properties)
Bytes dt_string(uint32_t size) {  // fixed-length, null-terminated ASCII
  Bytes b{0x13, 0x00, 0x00, 0x00};
  put(b, size, 4);
  return b;
}
Bytes dt_int(uint32_t size) {  // signed little-endian
  Bytes b{0x10, 0x08, 0x00, 0x00};
  put(b, size, 4);
  put(b, 0, 2);          // bit offset
  put(b, 8 * size, 2);   // precision
  return b;
}
Bytes dt_float(uint32_t size) {  // IEEE little-endian, implied leading mantissa bit
  const bool dbl = size == 8;
  Bytes b{0x11, 0x20, (unsigned char)(dbl ? 63 : 31), 0x00};
  put(b, size, 4);
  put(b, 0, 2);
  put(b, 8 * size, 2);
  b.push_back(dbl ? 52 : 23);   // exponent location
  b.push_back(dbl ? 11 : 8);    // exponent size
  b.push_back(0);               // mantissa location
  b.push_back(dbl ? 52 : 23);   // mantissa size
  put(b, dbl ? 1023 : 127, 4);  // exponent bias
  return b;
}
Bytes dataspace(const std::vector<uint64_t>& dims) {  // version 2, simple
  Bytes b{0x02, (unsigned char)dims.size(), 0x00, 0x01};
  for (auto d : dims) put(b, d, 8);
  return b;
}

struct Message {
  unsigned char type, flags;
  Bytes body;
};

// The file image; objects are appended and addressed by their offset.
struct File {
  Bytes img = Bytes(48, 0);  // superblock, filled in last

  void align8() {
    while (img.size() % 8) img.push_back(0);
  }
  // version-2 object header, 4-byte chunk size, one chunk, lookup3 checksum
  uint64_t object(const std::vector<Message>& msgs) {
    align8();
    const uint64_t at = img.size();
    Bytes b{'O', 'H', 'D', 'R', 0x02, 0x02};
    uint32_t size = 0;
    for (auto& m : msgs) size += 4 + (uint32_t)m.body.size();
    put(b, size, 4);
    for (auto& m : msgs) {
      b.push_back(m.type);
      put(b, m.body.size(), 2);
      b.push_back(m.flags);
      put(b, m.body.data(), m.body.size());
    }
    put(b, lookup3(b.data(), b.size()), 4);
    put(img, b.data(), b.size());
    return at;
  }
  uint64_t dataset(const Bytes& dtype, const std::vector<uint64_t>& dims, const void* raw, size_t bytes) {
    std::vector<Message> m;
    m.push_back({0x01, 0x00, dataspace(dims)});
    m.push_back({0x03, 0x01, dtype});
    Bytes layout{0x03};
    if (bytes <= 1024) {  // compact: the data lives in the header
      m.push_back({0x05, 0x01, {0x03, 0x09}});
      layout.push_back(0x00);
      put(layout, bytes, 2);
      put(layout, raw, bytes);
    } else {              // contiguous, written just before the header
      align8();
      const uint64_t addr = img.size();
      put(img, raw, bytes);
      m.push_back({0x05, 0x01, {0x03, 0x0A}});
      layout.push_back(0x01);
      put(layout, addr, 8);
      put(layout, bytes, 8);
    }
    m.push_back({0x08, 0x00, layout});
    return object(m);
  }
};

Message attr_message(const char* name, const Bytes& dtype, const std::vector<uint64_t>& dims, const void* raw,
                     size_t bytes) {
  Bytes b{0x03, 0x00};
  const size_t nlen = std::strlen(name) + 1;
  Bytes ds = dims.empty() ? Bytes{0x02, 0x00, 0x00, 0x00} : dataspace(dims);
  put(b, nlen, 2);
  put(b, dtype.size(), 2);
  put(b, ds.size(), 2);
  b.push_back(0x00);  // ASCII name
  put(b, name, nlen);
  put(b, dtype.data(), dtype.size());
  put(b, ds.data(), ds.size());
  put(b, raw, bytes);
  return {0x0C, 0x00, b};
}

Message attr_string(const char* name, const std::string& value, uint32_t width) {
  std::vector<char> s(width, 0);
  std::memcpy(s.data(), value.data(), std::min<size_t>(value.size(), width - 1));
  return attr_message(name, dt_string(width), {}, s.data(), width);
}

// A CGNS node before serialisation.
struct Node {
  std::string name, label, type;  // type: MT, I4, R4, R8, C1
  std::vector<uint64_t> dims;     // CGNS (Fortran) order
  Bytes data;
  std::vector<Node> kids;
};

Node node(const std::string& name, const std::string& label, const std::string& type) {
  Node n;
  n.name = name;
  n.label = label;
  n.type = type;
  return n;
}
template <class T> Node node(const std::string& name, const std::string& label, const std::string& type,
                             const std::vector<uint64_t>& dims, const T* v, size_t count) {
  Node n = node(name, label, type);
  n.dims = dims;
  put(n.data, v, count * sizeof(T));
  return n;
}
Node node_text(const std::string& name, const std::string& label, const std::string& text) {
  return node(name, label, "C1", {text.size()}, text.data(), text.size());
}

Bytes dtype_of(const std::string& t) {
  if (t == "I4") return dt_int(4);
  if (t == "R4") return dt_float(4);
  if (t == "R8") return dt_float(8);
  return dt_int(1);  // C1: characters as 8-bit integers
}

// Group with links in creation order (as libcgns creates its groups).
std::vector<Message> group_messages(const std::vector<std::pair<std::string, uint64_t>>& links) {
  std::vector<Message> m;
  Bytes li{0x00, 0x03};
  put(li, links.size(), 8);  // maximum creation index
  put(li, UNDEF, 8);         // fractal heap (none: compact storage)
  put(li, UNDEF, 8);         // name index
  put(li, UNDEF, 8);         // creation-order index
  m.push_back({0x02, 0x00, li});
  m.push_back({0x0A, 0x01, {0x00, 0x00}});
  Bytes ai{0x00, 0x00};
  put(ai, UNDEF, 8);
  put(ai, UNDEF, 8);
  m.push_back({0x15, 0x04, ai});
  for (size_t i = 0; i < links.size(); i++) {
    Bytes l{0x01, 0x04};
    put(l, i, 8);
    l.push_back((unsigned char)links[i].first.size());
    put(l, links[i].first.data(), links[i].first.size());
    put(l, links[i].second, 8);
    m.push_back({0x06, 0x00, l});
  }
  return m;
}

uint64_t write_node(File& f, const Node& n) {
  std::vector<std::pair<std::string, uint64_t>> links;
  if (!n.data.empty()) {
    std::vector<uint64_t> cdims(n.dims.rbegin(), n.dims.rend());  // HDF5 dims are C order
    links.push_back({" data", f.dataset(dtype_of(n.type), cdims, n.data.data(), n.data.size())});
  }
  for (auto& k : n.kids) links.push_back({k.name, write_node(f, k)});
  auto m = group_messages(links);
  m.push_back(attr_string("name", n.name, 33));
  m.push_back(attr_string("label", n.label, 33));
  m.push_back(attr_string("type", n.type, 3));
  const int32_t flags = 1;
  m.push_back(attr_message("flags", dt_int(4), {1}, &flags, 4));
  return f.object(m);
}

}  // namespace

std::string write_cgns(const std::string& path, const CgnsSolution& s) {
  const int dim = s.dim;
  if (dim < 1 || dim > 3) return "dimension must be 1, 2 or 3";
  const int nvar = dim + 2;
  if ((int)s.fields.size() != nvar) return "expected dim + 2 fields";
  uint64_t nver = 1, ncel = 1;
  for (int d = 0; d < dim; d++) {
    if (s.vertices[d] < 2) return "a direction has fewer than 2 vertices";
    nver *= s.vertices[d];
    ncel *= s.vertices[d] - 1;
  }
  for (auto& v : s.fields)
    if (v.size() != ncel) return "field size does not match the zone";

  const float version = 4.0f;
  Node lib = node("CGNSLibraryVersion", "CGNSLibraryVersion_t", "R4", {1}, &version, 1);
  const int32_t base_dims[2] = {dim, dim};
  Node base = node("Base", "CGNSBase_t", "I4", {2}, base_dims, 2);
  std::vector<int32_t> zsize(3 * dim);
  for (int d = 0; d < dim; d++) {
    zsize[d] = s.vertices[d];
    zsize[dim + d] = s.vertices[d] - 1;
    zsize[2 * dim + d] = 0;
  }
  Node zone = node("Zone  1", "Zone_t", "I4", {(uint64_t)dim, 3}, zsize.data(), zsize.size());
  zone.kids.push_back(node_text("ZoneType", "ZoneType_t", "Structured"));

  std::vector<uint64_t> vdims, cdims;
  for (int d = 0; d < dim; d++) {
    vdims.push_back(s.vertices[d]);
    cdims.push_back(s.vertices[d] - 1);
  }
  Node coords = node("GridCoordinates", "GridCoordinates_t", "MT");
  const char* cname[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
  std::vector<double> x(nver);
  for (int d = 0; d < dim; d++) {
    // Initializer::initBlock (zone_FD/src/src/initializer.cpp:612-627): lo_d + index * h
    for (uint64_t c = 0; c < nver; c++) {
      uint64_t idx = c;
      for (int e = 0; e < d; e++) idx /= s.vertices[e];
      idx %= s.vertices[d];
      x[c] = s.lo[d] + idx * s.h;
    }
    coords.kids.push_back(node(cname[d], "DataArray_t", "R8", vdims, x.data(), nver));
  }
  zone.kids.push_back(coords);

  Node sol = node("flowSolution", "FlowSolution_t", "MT");
  sol.kids.push_back(node_text("GridLocation", "GridLocation_t", "CellCenter"));
  const char* vel[3] = {"XVelocity", "YVelocity", "ZVelocity"};
  sol.kids.push_back(node("Density", "DataArray_t", "R8", cdims, s.fields[0].data(), ncel));
  for (int d = 0; d < dim; d++) sol.kids.push_back(node(vel[d], "DataArray_t", "R8", cdims, s.fields[1 + d].data(), ncel));
  sol.kids.push_back(node("Pressure", "DataArray_t", "R8", cdims, s.fields[dim + 1].data(), ncel));
  zone.kids.push_back(sol);
  base.kids.push_back(zone);

  File f;
  // the root group: the MotherNode of ADF-on-HDF5, with its format strings
  std::vector<std::pair<std::string, uint64_t>> links;
  const char fmt[] = "IEEE_LITTLE_32";
  links.push_back({" format", f.dataset(dt_int(1), {sizeof fmt}, fmt, sizeof fmt)});
  char h5v[33] = {0};
  std::snprintf(h5v, sizeof h5v, "HDF5 Version 1.10.0");
  links.push_back({" hdf5version", f.dataset(dt_int(1), {sizeof h5v}, h5v, sizeof h5v)});
  links.push_back({lib.name, write_node(f, lib)});
  links.push_back({base.name, write_node(f, base)});
  auto m = group_messages(links);
  m.push_back(attr_string("name", "HDF5 MotherNode", 33));
  m.push_back(attr_string("label", "Root Node of HDF5 File", 33));
  m.push_back(attr_string("type", "MT", 3));
  const uint64_t root = f.object(m);

  // version-2 superblock
  Bytes sb{0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n', 0x02, 0x08, 0x08, 0x00};
  put(sb, 0, 8);                 // base address
  put(sb, UNDEF, 8);             // superblock extension
  put(sb, f.img.size(), 8);      // end of file
  put(sb, root, 8);
  put(sb, lookup3(sb.data(), sb.size()), 4);
  std::memcpy(f.img.data(), sb.data(), sb.size());

  std::unique_ptr<std::FILE, int (*)(std::FILE*)> fp(std::fopen(path.c_str(), "wb"), std::fclose);
  if (!fp) return "cannot open " + path;
  if (std::fwrite(f.img.data(), 1, f.img.size(), fp.get()) != f.img.size()) return "short write to " + path;
  return "";
}

}  // namespace cfdhost
