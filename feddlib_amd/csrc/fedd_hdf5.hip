// EpetraExt-HDF5 multivector reader (host only).
//
// FEDDLib stores reference solutions with EpetraExt::HDF5 (HDF5Export / HDF5Import,
// feddlib/core/General/HDF5Import_def.hpp:22-70): one group per variable holding the
// datasets Values (NumVectors x GlobalLength, FP64), GlobalLength and NumVectors (int32)
// and a __type__ string.  EpetraExt writes through a linear map, so position k of Values
// is global id k.  No HDF5 library is linked: this file reads the subset of the HDF5
// file format that such files use, from the published format specification:
//   - superblock version 0 or 1;
//   - old-style groups: symbol-table message, version-1 B-tree of group nodes,
//     symbol-table nodes and a local heap for the names;
//   - version-1 object headers, with continuation blocks;
//   - dataspace version 1 or 2; fixed-point and IEEE floating-point datatypes,
//     little-endian; data layout version 1-3, contiguous or compact storage.
// Filters, chunked storage and version-2 object headers are refused with an error.
// Every read is bounds-checked against the file size.

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fedd_internal.h"

namespace fedd {

namespace {

struct H5Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Dataset {
  std::vector<uint64_t> dims;
  int tclass = -1;  // 0 fixed point, 1 floating point
  int tsize = 0;
  bool tsigned = false;
  bool big_endian = false;
  int layout = -1;  // 0 compact, 1 contiguous
  uint64_t addr = 0, size = 0;
  std::vector<uint8_t> compact;
  bool filtered = false;
};

class H5File {
 public:
  explicit H5File(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) throw H5Error(std::string("cannot open ") + path);
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0 || n > (1L << 31)) {
      std::fclose(f);
      throw H5Error(std::string("unreadable or too large: ") + path);
    }
    d_.resize((size_t)n);
    const size_t got = n ? std::fread(d_.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    if (got != (size_t)n) throw H5Error(std::string("short read: ") + path);
    superblock();
  }

  // Object header address of `name` in the group whose object header is at `group`.
  bool find(uint64_t group, const std::string& name, uint64_t* ohdr) const {
    uint64_t btree = 0, heap = 0;
    bool symtab = false;
    messages(group, [&](int type, uint64_t o, uint64_t sz) {
      if (type == 0x11) {
        need(o, 2 * so_, sz);
        btree = u(o, so_);
        heap = u(o + so_, so_);
        symtab = true;
      }
    });
    if (!symtab) throw H5Error("group without a symbol table (only old-style groups are read)");
    const uint64_t names = heap_data(heap);
    bool found = false;
    walk_btree(btree, 0, [&](uint64_t name_off, uint64_t oh) {
      if (!found && cstr(names, name_off) == name) {
        *ohdr = oh;
        found = true;
      }
    });
    return found;
  }

  uint64_t root() const { return root_; }

  Dataset dataset(uint64_t oh) const {
    Dataset ds;
    bool space = false;
    messages(oh, [&](int type, uint64_t o, uint64_t sz) {
      if (type == 0x01) {  // dataspace
        need(o, 8, sz);
        const int ver = d_[o], rank = d_[o + 1], flags = d_[o + 2];
        if (ver != 1 && ver != 2) throw H5Error("unsupported dataspace version");
        if (rank > 32) throw H5Error("dataspace rank out of range");
        const uint64_t p = o + (ver == 1 ? 8 : 4);
        need(p, (uint64_t)rank * sl_, sz - (p - o));
        for (int k = 0; k < rank; ++k) ds.dims.push_back(u(p + (uint64_t)k * sl_, sl_));
        (void)flags;
        space = true;
      } else if (type == 0x03) {  // datatype
        need(o, 8, sz);
        ds.tclass = d_[o] & 0x0f;
        const int bits = d_[o + 1];
        ds.tsize = (int)u(o + 4, 4);
        ds.big_endian = bits & 1;
        ds.tsigned = ds.tclass == 0 && (bits & 8);
      } else if (type == 0x08) {  // data layout
        need(o, 2, sz);
        const int ver = d_[o];
        if (ver == 3 || ver == 4) {
          ds.layout = d_[o + 1];
          if (ds.layout == 0) {
            need(o + 2, 2, sz - 2);
            const uint64_t n = u(o + 2, 2);
            need(o + 4, n, sz - 4);
            ds.compact.assign(d_.begin() + (long)(o + 4), d_.begin() + (long)(o + 4 + n));
          } else if (ds.layout == 1) {
            need(o + 2, so_ + sl_, sz - 2);
            ds.addr = u(o + 2, so_);
            ds.size = u(o + 2 + so_, sl_);
          }
        } else if (ver == 1 || ver == 2) {
          need(o, 8, sz);
          const int rank = d_[o + 1];
          ds.layout = d_[o + 2];
          uint64_t p = o + 8;
          if (ds.layout != 0) {
            need(p, so_, sz - 8);
            ds.addr = u(p, so_);
            p += so_;
          }
          p += 4ull * rank;
          if (p - o > sz) throw H5Error("truncated data layout message");
          if (ds.layout == 0) {
            need(p, 4, sz - (p - o));
            const uint64_t n = u(p, 4);
            need(p + 4, n, sz - (p + 4 - o));
            ds.compact.assign(d_.begin() + (long)(p + 4), d_.begin() + (long)(p + 4 + n));
          } else {
            ds.size = 0;  // computed from the dataspace below
          }
        } else {
          throw H5Error("unsupported data layout version");
        }
      } else if (type == 0x0b) {
        ds.filtered = true;
      }
    });
    if (!space || ds.tclass < 0 || ds.layout < 0) throw H5Error("object is not a dataset");
    if (ds.filtered) throw H5Error("filtered (compressed) datasets are not supported");
    if (ds.layout != 0 && ds.layout != 1) throw H5Error("only contiguous or compact datasets are supported");
    return ds;
  }

  // Dataset values converted to double.
  std::vector<double> read(const Dataset& ds) const {
    uint64_t n = 1;
    for (uint64_t x : ds.dims) {
      if (x && n > (1ull << 40) / x) throw H5Error("dataset too large");
      n *= x;
    }
    if (ds.big_endian) throw H5Error("big-endian data is not supported");
    const int w = ds.tsize;
    if (ds.tclass == 1 ? (w != 4 && w != 8) : ds.tclass == 0 ? (w != 1 && w != 2 && w != 4 && w != 8) : true)
      throw H5Error("dataset type is neither a plain integer nor an IEEE float");
    const uint64_t bytes = n * (uint64_t)w;
    const uint8_t* src;
    if (ds.layout == 0) {
      if (ds.compact.size() < bytes) throw H5Error("compact dataset shorter than its dataspace");
      src = ds.compact.data();
    } else {
      if (ds.addr == ~0ull) throw H5Error("dataset has no storage");
      if (ds.size && ds.size < bytes) throw H5Error("contiguous dataset shorter than its dataspace");
      const uint64_t o = addr(ds.addr);
      need(o, bytes, d_.size());
      src = d_.data() + o;
    }
    std::vector<double> out(n);
    for (uint64_t k = 0; k < n; ++k) {
      const uint8_t* p = src + k * (uint64_t)w;
      if (ds.tclass == 1) {
        if (w == 8) {
          double v;
          std::memcpy(&v, p, 8);
          out[k] = v;
        } else {
          float v;
          std::memcpy(&v, p, 4);
          out[k] = v;
        }
      } else {
        uint64_t v = 0;
        for (int b = 0; b < w; ++b) v |= (uint64_t)p[b] << (8 * b);
        if (ds.tsigned && w < 8 && (v >> (8 * w - 1)) & 1) v |= ~0ull << (8 * w);
        out[k] = ds.tsigned ? (double)(int64_t)v : (double)v;
      }
    }
    return out;
  }

 private:
  std::vector<uint8_t> d_;
  int so_ = 8, sl_ = 8;
  uint64_t base_ = 0, root_ = 0;

  void need(uint64_t o, uint64_t n, uint64_t avail) const {
    if (n > avail || o > d_.size() || n > d_.size() - o) throw H5Error("truncated or corrupt HDF5 file");
  }
  uint64_t u(uint64_t o, int n) const {
    need(o, (uint64_t)n, d_.size());
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v |= (uint64_t)d_[o + i] << (8 * i);
    return v;
  }
  uint64_t addr(uint64_t a) const {  // relative address -> file offset
    if (a == ~0ull || a > d_.size() || base_ > d_.size() - a) throw H5Error("address outside the file");
    return a + base_;
  }

  void superblock() {
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    uint64_t at = ~0ull;
    for (uint64_t o = 0; o + 8 <= d_.size(); o = o ? 2 * o : 512)
      if (std::memcmp(d_.data() + o, sig, 8) == 0) {
        at = o;
        break;
      }
    if (at == ~0ull) throw H5Error("not an HDF5 file");
    need(at, 24, d_.size());
    const int ver = d_[at + 8];
    if (ver != 0 && ver != 1) throw H5Error("unsupported HDF5 superblock version (0 or 1 are read)");
    so_ = d_[at + 13];
    sl_ = d_[at + 14];
    if ((so_ != 4 && so_ != 8) || (sl_ != 4 && sl_ != 8)) throw H5Error("unsupported offset/length size");
    uint64_t p = at + 24 + (ver == 1 ? 4 : 0);
    base_ = u(p, so_);
    if (base_ != at && base_ != 0) throw H5Error("unsupported base address");
    const uint64_t eof = u(p + 2ull * so_, so_);  // base, free-space, end-of-file, driver-info
    if (eof > d_.size() - base_) throw H5Error("truncated HDF5 file (end-of-file address beyond the file)");
    p += 4ull * so_;
    root_ = addr(u(p + so_, so_));  // symbol table entry: name offset, object header
  }

  uint64_t heap_data(uint64_t h) const {
    const uint64_t o = addr(h);
    need(o, 8 + 2 * sl_ + so_, d_.size());
    if (std::memcmp(&d_[o], "HEAP", 4) != 0) throw H5Error("bad local heap");
    return addr(u(o + 8 + 2 * sl_, so_));
  }

  std::string cstr(uint64_t base, uint64_t off) const {
    std::string s;
    for (uint64_t o = base + off; o < d_.size() && d_[o]; ++o) {
      s.push_back((char)d_[o]);
      if (s.size() > 4096) throw H5Error("name too long");
    }
    return s;
  }

  template <class F>
  void walk_btree(uint64_t node, int depth, F&& f) const {
    if (depth > 32) throw H5Error("B-tree too deep");
    const uint64_t o = addr(node);
    need(o, 8 + 2 * so_, d_.size());
    if (std::memcmp(&d_[o], "TREE", 4) != 0 || d_[o + 4] != 0) throw H5Error("bad group B-tree node");
    const int level = d_[o + 5];
    const uint64_t n = u(o + 6, 2);
    uint64_t p = o + 8 + 2 * so_;
    for (uint64_t k = 0; k < n; ++k) {
      p += sl_;  // key
      const uint64_t child = u(p, so_);
      p += so_;
      if (level > 0) {
        walk_btree(child, depth + 1, f);
        continue;
      }
      const uint64_t s = addr(child);
      need(s, 8, d_.size());
      if (std::memcmp(&d_[s], "SNOD", 4) != 0) throw H5Error("bad symbol table node");
      const uint64_t ns = u(s + 6, 2);
      const uint64_t esz = 2 * so_ + 24;
      for (uint64_t e = 0; e < ns; ++e) {
        const uint64_t q = s + 8 + e * esz;
        f(u(q, so_), addr(u(q + so_, so_)));
      }
    }
  }

  // Calls f(type, body offset, body size) for every message of a version-1 object header.
  template <class F>
  void messages(uint64_t oh, F&& f) const {
    need(oh, 16, d_.size());
    if (d_[oh] != 1) throw H5Error("unsupported object header version (version 1 is read)");
    const uint64_t nmsg = u(oh + 2, 2);
    std::vector<std::pair<uint64_t, uint64_t>> blocks{{oh + 16, u(oh + 8, 4)}};
    uint64_t seen = 0;
    for (size_t b = 0; b < blocks.size() && seen < nmsg; ++b) {
      if (b > 256) throw H5Error("too many object header blocks");
      uint64_t o = blocks[b].first;
      const uint64_t end = o + blocks[b].second;
      need(blocks[b].first, blocks[b].second, d_.size());
      while (o + 8 <= end && seen < nmsg) {
        const int type = (int)u(o, 2);
        const uint64_t sz = u(o + 2, 2);
        need(o + 8, sz, end - (o + 8));
        if (type == 0x10) {
          need(o + 8, so_ + sl_, sz);
          blocks.emplace_back(addr(u(o + 8, so_)), u(o + 8 + so_, sl_));
        } else {
          f(type, o + 8, sz);
        }
        ++seen;
        o += 8 + sz;
      }
    }
  }
};

thread_local std::string g_free_err;

}  // namespace

std::string& free_error() { return g_free_err; }

}  // namespace fedd

extern "C" {

int fedd_hdf5_read_vector(const char* path, const char* name, double* out, int64_t cap,
                          int64_t* global_length, int* num_vectors) {
  using namespace fedd;
  if (!path || !name || cap < 0 || (cap > 0 && !out)) {
    g_free_err = "fedd_hdf5_read_vector: bad argument";
    return FEDD_ERR_ARG;
  }
  try {
    H5File f(path);
    uint64_t grp = 0;
    if (!f.find(f.root(), name, &grp))
      throw H5Error(std::string("Requested varName: ") + name + " not contained in hdf file " + path + ".");
    uint64_t o_vals = 0, o_len = 0, o_nv = 0;
    if (!f.find(grp, "Values", &o_vals) || !f.find(grp, "GlobalLength", &o_len) || !f.find(grp, "NumVectors", &o_nv))
      throw H5Error(std::string("group ") + name + " is not an EpetraExt multivector");
    const std::vector<double> len = f.read(f.dataset(o_len)), nv = f.read(f.dataset(o_nv));
    if (len.size() != 1 || nv.size() != 1 || !(len[0] >= 0) || !(nv[0] >= 1))
      throw H5Error("bad GlobalLength / NumVectors");
    const int64_t n = (int64_t)len[0];
    const int k = (int)nv[0];
    const Dataset ds = f.dataset(o_vals);
    const std::vector<double> v = f.read(ds);
    if ((int64_t)v.size() != n * k) throw H5Error("Values does not hold NumVectors x GlobalLength entries");
    if (global_length) *global_length = n;
    if (num_vectors) *num_vectors = k;
    if (out) {
      if (cap < n) throw H5Error("output buffer too small");
      // Values is NumVectors x GlobalLength (EpetraExt's untransposed layout): vector 0
      // is the first row; a GlobalLength x NumVectors dataset holds it in column 0.
      const bool transposed = ds.dims.size() == 2 && (int64_t)ds.dims[0] == n && (int64_t)ds.dims[1] == k && k > 1;
      for (int64_t i = 0; i < n; ++i) out[i] = transposed ? v[(size_t)(i * k)] : v[(size_t)i];
    }
    return FEDD_OK;
  } catch (const std::exception& e) {
    g_free_err = e.what();
    return FEDD_ERR_ARG;
  }
}

}  // extern "C"
