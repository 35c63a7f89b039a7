// writers.cpp — result output in the reference's file formats:
//   legacy-ASCII VTK        write_vtk.hpp:18-252
//   one CSV per member      write_csv.hpp:18-181
//   Forst struct of chunks  abstract/forst.hpp:51-615, forst_chunk.hpp:26-69
// Host code over the public C ABI: the fields come back from the device through
// lbm_get_*, and the text is formatted on all host cores.  Numbers go through
// std::to_chars(general, 6), which is specified as printf's %.6g — the
// conversion the reference's `ostream << value` performs with default flags —
// so the bytes match the reference writers' for the same values.
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lbm_mirror_amd/lbm.h"

lbm_status lbmx_report(lbm_status s, const char* msg);

namespace {

enum class Kind { Vector, Scalar, Flag, Pops };

struct Field {
  std::string key;
  Kind kind = Kind::Scalar;
  int comps = 1;                // values per cell
  int ghost = 0;                // chunk ghost ring (Forst extents)
  const double* v = nullptr;    // SoA: v[k*N + cell]
  const uint8_t* fl = nullptr;  // flag chunk
};

struct Shape {
  int32_t lattice = 0, storage = 0, D = 2, Q = 9;
  uint64_t n[3] = {1, 1, 1};
  uint64_t N = 0;
};

lbm_status report(lbm_status s, const std::string& m) { return lbmx_report(s, m.c_str()); }

lbm_status make_shape(int32_t lattice, int32_t storage, const uint64_t n[3], Shape& s) {
  if (lattice < LBM_D2Q9 || lattice > LBM_D3Q27) return report(LBM_ERR_INVALID_ARG, "unknown lattice");
  if (storage < LBM_F64 || storage > LBM_BF16S) return report(LBM_ERR_INVALID_ARG, "unknown storage");
  s.lattice = lattice;
  s.storage = storage;
  s.D = lattice == LBM_D2Q9 ? 2 : 3;
  s.Q = lattice == LBM_D2Q9 ? 9 : (lattice == LBM_D3Q19 ? 19 : 27);
  s.N = 1;
  for (int d = 0; d < 3; ++d) {
    s.n[d] = d < s.D ? n[d] : 1;
    if (s.n[d] == 0) return report(LBM_ERR_INVALID_ARG, "empty extent");
    s.N *= s.n[d];
  }
  return LBM_OK;
}

// Points each parsed member at its buffer in the field set.
lbm_status bind(const lbm_field_set& set, std::vector<Field>& fs) {
  for (Field& f : fs) {
    if (f.key == "velocity") f.v = set.velocity;
    else if (f.key == "density") f.v = set.density;
    else if (f.key == "force") f.v = set.force;
    else if (f.key == "porosity") f.v = set.porosity;
    else if (f.key == "dfs") f.v = set.dfs;
    else if (f.key == "info") f.fl = set.info;
    if (!f.v && !f.fl) return report(LBM_ERR_INVALID_ARG, "no buffer for member '" + f.key + "'");
  }
  return LBM_OK;
}

// Parses "velocity,density,..." into typed (still empty) fields.
lbm_status parse_members(const char* members, const Shape& s, std::vector<Field>& out) {
  if (!members) return report(LBM_ERR_INVALID_ARG, "null member list");
  std::string all(members);
  size_t p = 0;
  while (p <= all.size()) {
    size_t q = all.find(',', p);
    if (q == std::string::npos) q = all.size();
    std::string k = all.substr(p, q - p);
    k.erase(0, k.find_first_not_of(" \t"));
    k.erase(k.find_last_not_of(" \t") + 1);
    if (!k.empty()) {
      Field f;
      f.key = k;
      if (k == "velocity" || k == "force") {
        f.kind = Kind::Vector;
        f.comps = s.D;
      } else if (k == "density" || k == "porosity") {
        f.kind = Kind::Scalar;
      } else if (k == "info") {
        f.kind = Kind::Flag;
      } else if (k == "dfs") {
        f.kind = Kind::Pops;
        f.comps = s.Q;
        f.ghost = 1;
      } else {
        return report(LBM_ERR_NOT_FOUND, "unknown member '" + k +
                                             "' (velocity, density, force, porosity, info, dfs)");
      }
      for (const Field& g : out)
        if (g.key == k) return report(LBM_ERR_INVALID_ARG, "member '" + k + "' listed twice");
      out.push_back(std::move(f));
    }
    p = q + 1;
  }
  return LBM_OK;
}

// Reads the listed members of a context back to host buffers it owns.
struct Fetched {
  lbm_field_set set{};
  std::vector<double> rho, u, force, por, dfs;
  std::vector<uint8_t> info;
};

lbm_status fetch(lbm_ctx* ctx, const char* members, Fetched& out) {
  if (!ctx) return report(LBM_ERR_INVALID_ARG, "null ctx");
  lbm_field_set& set = out.set;
  lbm_status st = lbm_ctx_info(ctx, &set.lattice, &set.storage, set.extent);
  Shape s;
  std::vector<Field> fs;
  if (!st) st = make_shape(set.lattice, set.storage, set.extent, s);
  if (!st) st = parse_members(members, s, fs);
  for (const Field& f : fs) {
    if (st) break;
    if (f.key == "velocity" || f.key == "density") {
      if (out.rho.empty()) {
        out.rho.resize(s.N);
        out.u.resize(size_t(s.D) * s.N);
        st = lbm_get_macros(ctx, out.rho.data(), out.u.data());
        set.density = out.rho.data();
        set.velocity = out.u.data();
      }
    } else if (f.key == "force") {
      out.force.resize(size_t(s.D) * s.N);
      st = lbm_get_force_field(ctx, out.force.data());
      set.force = out.force.data();
    } else if (f.key == "porosity") {
      out.por.resize(s.N);
      st = lbm_get_porosity(ctx, out.por.data());
      set.porosity = out.por.data();
    } else if (f.key == "info") {
      out.info.resize(s.N);
      st = lbm_get_flags(ctx, out.info.data());
      set.info = out.info.data();
    } else if (f.key == "dfs") {
      out.dfs.resize(size_t(s.Q) * s.N);
      st = lbm_get_populations(ctx, out.dfs.data());
      set.dfs = out.dfs.data();
    }
  }
  return st;
}

// ---- text ---------------------------------------------------------------------

inline char* put_num(char* p, double x) {
  // 32 bytes hold any %.6g rendering ("-1.23457e-308" is 13)
  return std::to_chars(p, p + 32, x, std::chars_format::general, 6).ptr;
}

inline char* put_uint(char* p, unsigned v) { return std::to_chars(p, p + 8, v).ptr; }

// One line per cell: the values separated by `sep`, then `pad` copies of " 0"
// (VTK pads vectors to three components), then a newline.  `narrow` rounds
// to float first: the value a float field of the reference holds.
char* put_cell(char* p, const Field& f, uint64_t N, uint64_t c, char sep, int pad, bool narrow) {
  if (f.kind == Kind::Flag) {
    p = put_uint(p, f.fl[c]);
  } else {
    for (int k = 0; k < f.comps; ++k) {
      if (k) *p++ = sep;
      const double x = f.v[size_t(k) * N + c];
      p = put_num(p, narrow ? double(float(x)) : x);
    }
  }
  for (int k = 0; k < pad; ++k) {
    *p++ = ' ';
    *p++ = '0';
  }
  *p++ = '\n';
  return p;
}

unsigned worker_count() {
  unsigned h = std::thread::hardware_concurrency();
  return std::max(1u, std::min(h ? h : 1u, 16u));
}

// Formats the cells of one field on all workers (contiguous cell ranges, one
// buffer each) and appends the pieces to `out` in cell order.
lbm_status emit_cells(std::FILE* out, const Field& f, uint64_t N, char sep, int pad, bool narrow) {
  const unsigned W = unsigned(std::min<uint64_t>(worker_count(), (N + 4095) / 4096));
  const size_t per_cell = size_t(f.comps) * 32 + size_t(pad) * 2 + 2;
  std::vector<std::vector<char>> bufs(W);
  std::vector<size_t> len(W, 0);
  auto work = [&](unsigned w) {
    const uint64_t c0 = N * w / W, c1 = N * (w + 1) / W;
    bufs[w].resize((c1 - c0) * per_cell);
    char* p = bufs[w].data();
    for (uint64_t c = c0; c < c1; ++c) p = put_cell(p, f, N, c, sep, pad, narrow);
    len[w] = size_t(p - bufs[w].data());
  };
  std::vector<std::thread> th;
  for (unsigned w = 1; w < W; ++w) th.emplace_back(work, w);
  if (W) work(0);
  for (auto& t : th) t.join();
  for (unsigned w = 0; w < W; ++w)
    if (len[w] && std::fwrite(bufs[w].data(), 1, len[w], out) != len[w])
      return report(LBM_ERR_CRITICAL, "write failed");
  return LBM_OK;
}

struct File {
  std::FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

lbm_status open_in(const std::filesystem::path& dir, const std::string& name, File& file) {
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) return report(LBM_ERR_CRITICAL, "could not create directory " + dir.string());
  file.f = std::fopen((dir / name).c_str(), "wb");
  if (!file.f) return report(LBM_ERR_CRITICAL, "could not open " + (dir / name).string());
  return LBM_OK;
}

lbm_status close_file(File& file) {
  const bool bad = std::ferror(file.f) != 0;
  const bool fail = std::fclose(file.f) != 0;
  file.f = nullptr;
  return (bad || fail) ? report(LBM_ERR_CRITICAL, "write failed") : LBM_OK;
}

lbm_status prepare(const lbm_field_set* set, const char* members, Shape& s, std::vector<Field>& fs) {
  if (!set) return report(LBM_ERR_INVALID_ARG, "null field set");
  lbm_status st = make_shape(set->lattice, set->storage, set->extent, s);
  if (!st) st = parse_members(members, s, fs);
  if (!st) st = bind(*set, fs);
  return st;
}

// ---- Forst --------------------------------------------------------------------

size_t elem_bytes(const Field& f, const Shape& s) {
  return f.kind == Kind::Flag ? 1 : (s.storage == LBM_F64 ? 8 : 4);
}

uint64_t chunk_cells(const Field& f, const Shape& s) {
  uint64_t c = 1;
  for (int d = 0; d < s.D; ++d) c *= s.n[d] + 2 * uint64_t(f.ghost);
  return c;
}

uint64_t chunk_bytes(const Field& f, const Shape& s) {
  return chunk_cells(f, s) * uint64_t(f.comps) * elem_bytes(f, s);
}

// Runs fn(r0, r1) over the rows (y, z pairs) of the lattice on all workers.
template <typename Fn>
void for_rows(const Shape& s, Fn fn) {
  const uint64_t rows = s.n[1] * s.n[2];
  const unsigned W = unsigned(std::min<uint64_t>(worker_count(), std::max<uint64_t>(1, rows / 16)));
  std::vector<std::thread> th;
  for (unsigned w = 1; w < W; ++w) th.emplace_back(fn, rows * w / W, rows * (w + 1) / W);
  fn(uint64_t(0), rows / W);
  for (auto& t : th) t.join();
}

// Interior row r = y + ny*z starts at this stored (ghost-widened) cell.
uint64_t stored_row(const Field& f, const Shape& s, uint64_t r) {
  const uint64_t g = uint64_t(f.ghost);
  const uint64_t y = r % s.n[1], z = r / s.n[1];
  const uint64_t sx = s.n[0] + 2 * g, sy = s.n[1] + 2 * g;
  return ((s.D == 3 ? z + g : 0) * sy + (y + g)) * sx + g;
}

// SoA doubles -> the chunk's stored values (cell-major, components inner).
template <typename T>
void pack_chunk(const Field& f, const Shape& s, uint8_t* out) {
  T* o = reinterpret_cast<T*>(out);
  const uint64_t N = s.N, nx = s.n[0], K = uint64_t(f.comps);
  for_rows(s, [&](uint64_t r0, uint64_t r1) {
    for (uint64_t r = r0; r < r1; ++r) {
      T* dst = o + stored_row(f, s, r) * K;
      const uint64_t c0 = r * nx;
      for (uint64_t k = 0; k < K; ++k) {
        const double* src = f.v + k * N + c0;
        for (uint64_t x = 0; x < nx; ++x) dst[x * K + k] = T(src[x]);
      }
    }
  });
}

template <typename T>
void unpack_chunk(const Field& f, const Shape& s, const uint8_t* in, double* out) {
  const T* src = reinterpret_cast<const T*>(in);
  const uint64_t N = s.N, nx = s.n[0], K = uint64_t(f.comps);
  for_rows(s, [&](uint64_t r0, uint64_t r1) {
    for (uint64_t r = r0; r < r1; ++r) {
      const T* row = src + stored_row(f, s, r) * K;
      const uint64_t c0 = r * nx;
      for (uint64_t k = 0; k < K; ++k) {
        double* dst = out + k * N + c0;
        for (uint64_t x = 0; x < nx; ++x) dst[x] = double(row[x * K + k]);
      }
    }
  });
}

}  // namespace

extern "C" {

lbm_status lbm_write_vtk_fields(const lbm_field_set* set, const char* out_dir, const char* file_name,
                                uint64_t t, const char* members) {
  if (!out_dir || !file_name) return report(LBM_ERR_INVALID_ARG, "null path");
  Shape s;
  std::vector<Field> fs;
  lbm_status st = prepare(set, members, s, fs);
  if (st) return st;
  File file;
  st = open_in(std::filesystem::path(out_dir) / "vtk",
               std::string(file_name) + "_" + std::to_string(t) + ".vtk", file);
  if (st) return st;
  std::string head =
      "# vtk DataFile Version 3.0\nLBM File\nASCII\nDATASET STRUCTURED_POINTS\n"
      "SPACING 1.0 1.0 1.0\nORIGIN 0.0 0.0 0.0\nDIMENSIONS";
  for (int d = 0; d < 3; ++d) head += " " + std::to_string(s.n[d]);
  head += "\n";
  if (!fs.empty()) head += "\nPOINT_DATA " + std::to_string(s.N) + "\n\n";
  std::fputs(head.c_str(), file.f);
  for (const Field& f : fs) {
    const bool vec = f.kind == Kind::Vector || f.kind == Kind::Pops;
    const std::string h = vec ? "VECTORS " + f.key + " float\n"
                              : "SCALARS " + f.key + " float 1\nLOOKUP_TABLE default\n";
    std::fputs(h.c_str(), file.f);
    st = emit_cells(file.f, f, s.N, ' ', vec ? std::max(0, 3 - f.comps) : 0, s.storage != LBM_F64);
    if (st) return st;
    std::fputc('\n', file.f);
  }
  return close_file(file);
}

lbm_status lbm_write_csv_fields(const lbm_field_set* set, const char* out_dir, const char* file_name,
                                uint64_t t, const char* members) {
  if (!out_dir || !file_name) return report(LBM_ERR_INVALID_ARG, "null path");
  Shape s;
  std::vector<Field> fs;
  lbm_status st = prepare(set, members, s, fs);
  if (st) return st;
  const std::filesystem::path dir = std::filesystem::path(out_dir) / "csv";
  for (const Field& f : fs) {
    File file;
    st = open_in(dir, std::string(file_name) + "_" + f.key + "_" + std::to_string(t) + ".csv", file);
    if (!st) st = emit_cells(file.f, f, s.N, ',', 0, s.storage != LBM_F64);
    if (!st) st = close_file(file);
    if (st) return st;
  }
  return LBM_OK;
}

lbm_status lbm_write_forst_fields(const lbm_field_set* set, const char* path, const char* members) {
  if (!path) return report(LBM_ERR_INVALID_ARG, "null path");
  Shape s;
  std::vector<Field> fs;
  lbm_status st = prepare(set, members, s, fs);
  if (st) return st;
  for (const Field& f : fs)
    if (f.kind == Kind::Pops && s.storage != LBM_F64 && s.storage != LBM_F32)
      return report(LBM_ERR_NOT_SUPPORTED, "16-bit shifted populations have no Forst layout");
  File file;
  file.f = std::fopen(path, "wb");
  if (!file.f) return report(LBM_ERR_CRITICAL, std::string("could not open ") + path);
  for (const Field& f : fs) {
    std::vector<uint8_t> bytes(chunk_bytes(f, s), 0);
    if (f.kind == Kind::Flag)
      std::memcpy(bytes.data(), f.fl, s.N);
    else if (s.storage == LBM_F64)
      pack_chunk<double>(f, s, bytes.data());
    else
      pack_chunk<float>(f, s, bytes.data());
    if (std::fwrite(bytes.data(), 1, bytes.size(), file.f) != bytes.size())
      return report(LBM_ERR_CRITICAL, "write failed");
  }
  return close_file(file);
}

lbm_status lbm_read_forst_fields(lbm_field_set* set, const char* path, const char* members) {
  if (!path) return report(LBM_ERR_INVALID_ARG, "null path");
  Shape s;
  std::vector<Field> fs;
  lbm_status st = prepare(set, members, s, fs);
  if (st) return st;
  uint64_t need = 0;
  for (const Field& f : fs) {
    if (f.kind == Kind::Pops && s.storage != LBM_F64 && s.storage != LBM_F32)
      return report(LBM_ERR_NOT_SUPPORTED, "16-bit shifted populations have no Forst layout");
    need += chunk_bytes(f, s);
  }
  std::error_code ec;
  const uintmax_t have = std::filesystem::file_size(path, ec);
  if (ec) return report(LBM_ERR_NOT_FOUND, std::string("could not read ") + path);
  if (have < need) return report(LBM_ERR_INVALID_ARG, "Truncated Forst buffer");
  std::vector<uint8_t> bytes(need);
  {
    File file;
    file.f = std::fopen(path, "rb");
    if (!file.f || std::fread(bytes.data(), 1, need, file.f) != need)
      return report(LBM_ERR_CRITICAL, std::string("could not read ") + path);
  }
  uint64_t at = 0;
  for (const Field& f : fs) {
    const uint8_t* src = bytes.data() + at;
    if (f.kind == Kind::Flag)
      std::memcpy(const_cast<uint8_t*>(f.fl), src, s.N);
    else if (s.storage == LBM_F64)
      unpack_chunk<double>(f, s, src, const_cast<double*>(f.v));
    else
      unpack_chunk<float>(f, s, src, const_cast<double*>(f.v));
    at += chunk_bytes(f, s);
  }
  return LBM_OK;
}

lbm_status lbm_write_vtk(lbm_ctx* ctx, const char* out_dir, const char* file_name, uint64_t t,
                         const char* members) {
  Fetched h;
  lbm_status st = fetch(ctx, members, h);
  return st ? st : lbm_write_vtk_fields(&h.set, out_dir, file_name, t, members);
}

lbm_status lbm_write_csv(lbm_ctx* ctx, const char* out_dir, const char* file_name, uint64_t t,
                         const char* members) {
  Fetched h;
  lbm_status st = fetch(ctx, members, h);
  return st ? st : lbm_write_csv_fields(&h.set, out_dir, file_name, t, members);
}

lbm_status lbm_write_forst(lbm_ctx* ctx, const char* path, const char* members) {
  Fetched h;
  lbm_status st = fetch(ctx, members, h);
  return st ? st : lbm_write_forst_fields(&h.set, path, members);
}

lbm_status lbm_read_forst(lbm_ctx* ctx, const char* path, const char* members) {
  if (!ctx) return report(LBM_ERR_INVALID_ARG, "null ctx");
  lbm_field_set set{};
  lbm_status st = lbm_ctx_info(ctx, &set.lattice, &set.storage, set.extent);
  Shape s;
  std::vector<Field> fs;
  if (!st) st = make_shape(set.lattice, set.storage, set.extent, s);
  if (!st) st = parse_members(members, s, fs);
  if (st) return st;
  std::vector<double> rho(s.N), u(size_t(s.D) * s.N), force(size_t(s.D) * s.N), por(s.N),
      dfs(size_t(s.Q) * s.N);
  std::vector<uint8_t> info(s.N);
  set.density = rho.data();
  set.velocity = u.data();
  set.force = force.data();
  set.porosity = por.data();
  set.dfs = dfs.data();
  set.info = info.data();
  st = lbm_read_forst_fields(&set, path, members);
  if (st) return st;
  bool has[6] = {};
  for (const Field& f : fs) {
    if (f.key == "info") has[0] = true;
    else if (f.key == "dfs") has[1] = true;
    else if (f.key == "density") has[2] = true;
    else if (f.key == "velocity") has[3] = true;
    else if (f.key == "force") has[4] = true;
    else if (f.key == "porosity") has[5] = true;
  }
  // flags first: the other setters validate against them
  if (has[0]) st = lbm_set_flags(ctx, info.data());
  if (!st && has[1]) st = lbm_set_populations(ctx, dfs.data());
  if (!st && (has[2] || has[3]))
    st = lbm_set_macros(ctx, has[2] ? rho.data() : nullptr, has[3] ? u.data() : nullptr);
  if (!st && has[4]) st = lbm_set_force_field(ctx, force.data());
  if (!st && has[5]) st = lbm_set_porosity(ctx, por.data());
  return st;
}

}  // extern "C"
