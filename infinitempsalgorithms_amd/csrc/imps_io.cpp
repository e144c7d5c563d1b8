// State and model files: checkpoint and resume of a CanonicalIMPS (AL, AR, C, AC per site) and
// of the per-site MPO tensors, and the exchange of initial states between processes.
//
// Layout (little-endian):
//   bytes  0..7   magic "IMPSSTA1" (state) or "IMPSMPO1" (MPO)
//          8..11  uint32 format version (1)
//         12..15  uint32 0x01020304 (byte-order check)
//         16..27  int32 nsites, D (state) or w (MPO), d
//         28..31  int32 element type (1 = float64)
//         32..    payload, column-major float64: per site AL, AR (D d D each), C (D D), AC (D d D)
//                 for a state; W (w d w d) for an MPO
//   last 8 bytes  uint64 FNV-1a hash of the payload bytes
// Reads check the file's size against the size its header announces before allocating anything.
// A file is written to "<path>.tmp" and renamed onto <path> once complete, so an interrupted
// write leaves any earlier file at <path> intact.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "imps.h"
#include "imps_internal.h"

namespace {

const char kStateMagic[9] = "IMPSSTA1";
const char kMpoMagic[9] = "IMPSMPO1";
const uint32_t kVersion = 1;
const uint32_t kOrder = 0x01020304u;

struct Header {
  char magic[8];
  uint32_t version;
  uint32_t order;
  int32_t nsites, dim, d, dtype;
};
static_assert(sizeof(Header) == 32, "header layout");

struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= b[i];
      h *= 1099511628211ull;
    }
  }
};

int fail(int code, const std::string& msg) {
  imps::set_global_error(msg);
  return code;
}

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

// blocks: the payload as (pointer, count) pairs in file order
int write_file(const char* path, const char* magic, int nsites, int dim, int d,
               const std::vector<std::pair<const double*, size_t>>& blocks) {
  const std::string tmp = std::string(path) + ".tmp";
  Header h;
  std::memcpy(h.magic, magic, 8);
  h.version = kVersion;
  h.order = kOrder;
  h.nsites = nsites;
  h.dim = dim;
  h.d = d;
  h.dtype = 1;
  Fnv fnv;
  {
    File out;
    out.f = std::fopen(tmp.c_str(), "wb");
    if (!out.f) return fail(IMPS_E_IO, "cannot open " + tmp + " for writing");
    bool ok = std::fwrite(&h, sizeof h, 1, out.f) == 1;
    for (const auto& b : blocks) {
      if (!ok) break;
      ok = std::fwrite(b.first, sizeof(double), b.second, out.f) == b.second;
      fnv.add(b.first, b.second * sizeof(double));
    }
    ok = ok && std::fwrite(&fnv.h, sizeof fnv.h, 1, out.f) == 1;
    ok = (std::fflush(out.f) == 0) && ok;
    ok = (std::fclose(out.f) == 0) && ok;
    out.f = nullptr;
    if (!ok) {
      std::remove(tmp.c_str());
      return fail(IMPS_E_IO, "write to " + tmp + " failed");
    }
  }
  if (std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return fail(IMPS_E_IO, std::string("cannot rename ") + tmp + " to " + path);
  }
  return IMPS_OK;
}

// payload doubles of a file of this kind and these dimensions; false when it would not fit in
// 2^60 bytes (no real file, so the header is not trusted)
bool payload_doubles(const char* magic, int64_t nsites, int64_t dim, int64_t d, uint64_t& out) {
  const long double per = (magic == kStateMagic) ? 3.0L * dim * d * dim + (long double)dim * dim
                                                 : (long double)dim * d * dim * d;
  const long double total = per * nsites;
  if (!(total * 8.0L < 1152921504606846976.0L)) return false;
  out = (magic == kStateMagic) ? (uint64_t)nsites * (3ull * dim * d * dim + (uint64_t)dim * dim)
                               : (uint64_t)nsites * ((uint64_t)dim * d * dim * d);
  return true;
}

// checks the header and that the file holds exactly the payload it announces plus the hash
int read_header(FILE* f, const char* path, const char* magic, Header& h) {
  if (std::fread(&h, sizeof h, 1, f) != 1) return fail(IMPS_E_IO, std::string(path) + ": truncated header");
  if (std::memcmp(h.magic, magic, 8) != 0)
    return fail(IMPS_E_IO, std::string(path) + ": not a " + (magic == kStateMagic ? "state" : "MPO") + " file");
  if (h.version != kVersion) return fail(IMPS_E_IO, std::string(path) + ": unsupported format version");
  if (h.order != kOrder) return fail(IMPS_E_IO, std::string(path) + ": written with another byte order");
  if (h.dtype != 1) return fail(IMPS_E_IO, std::string(path) + ": element type is not float64");
  if (h.nsites <= 0 || h.dim <= 0 || h.d <= 0) return fail(IMPS_E_IO, std::string(path) + ": bad dimensions");
  uint64_t nd = 0;
  if (!payload_doubles(magic, h.nsites, h.dim, h.d, nd))
    return fail(IMPS_E_IO, std::string(path) + ": dimensions in the header are not plausible");
  const long here = std::ftell(f);
  if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return fail(IMPS_E_IO, std::string(path) + ": cannot seek");
  const long size = std::ftell(f);
  if (size < 0 || std::fseek(f, here, SEEK_SET) != 0) return fail(IMPS_E_IO, std::string(path) + ": cannot seek");
  if ((uint64_t)size < sizeof(Header) + 8 || (uint64_t)size - sizeof(Header) - 8 != nd * 8)
    return fail(IMPS_E_IO, std::string(path) + ": file size does not match its header (truncated or extended)");
  return IMPS_OK;
}

int info_file(const char* path, const char* magic, int* nsites, int* dim, int* d) {
  if (!path || !nsites || !dim || !d) return fail(IMPS_E_ARG, "null argument");
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return fail(IMPS_E_IO, std::string("cannot open ") + path);
  Header h;
  const int st = read_header(in.f, path, magic, h);
  if (st != IMPS_OK) return st;
  *nsites = h.nsites;
  *dim = h.dim;
  *d = h.d;
  return IMPS_OK;
}

// reads the payload into blocks after checking the header against (nsites, dim, d), the payload
// length and the hash; the destination is written only when every check passes
int read_file(const char* path, const char* magic, int nsites, int dim, int d,
              const std::vector<std::pair<double*, size_t>>& blocks) {
  File in;
  in.f = std::fopen(path, "rb");
  if (!in.f) return fail(IMPS_E_IO, std::string("cannot open ") + path);
  Header h;
  int st = read_header(in.f, path, magic, h);
  if (st != IMPS_OK) return st;
  if (h.nsites != nsites || h.dim != dim || h.d != d)
    return fail(IMPS_E_ARG, std::string(path) + ": dimensions differ from the requested ones");
  size_t total = 0;
  for (const auto& b : blocks) total += b.second;
  std::vector<double> buf(total);
  if (std::fread(buf.data(), sizeof(double), total, in.f) != total)
    return fail(IMPS_E_IO, std::string(path) + ": truncated payload");
  uint64_t stored = 0;
  if (std::fread(&stored, sizeof stored, 1, in.f) != 1) return fail(IMPS_E_IO, std::string(path) + ": truncated");
  char extra;
  if (std::fread(&extra, 1, 1, in.f) != 0) return fail(IMPS_E_IO, std::string(path) + ": trailing bytes");
  Fnv fnv;
  fnv.add(buf.data(), total * sizeof(double));
  if (fnv.h != stored) return fail(IMPS_E_IO, std::string(path) + ": checksum mismatch (corrupted file)");
  size_t off = 0;
  for (const auto& b : blocks) {
    std::memcpy(b.first, buf.data() + off, b.second * sizeof(double));
    off += b.second;
  }
  return IMPS_OK;
}

}  // namespace

// runs f, turning an exception (allocation failure) into an error code: nothing may unwind
// through the C interface
template <typename F>
int guarded(const char* who, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(IMPS_E_IO, std::string(who) + ": out of host memory");
  } catch (const std::exception& e) {
    return fail(IMPS_E_IO, std::string(who) + ": " + e.what());
  } catch (...) {
    return fail(IMPS_E_IO, std::string(who) + ": unexpected error");
  }
}

extern "C" {

int imps_write_state_file(const char* path, int nsites, int D, int d, const double* const* AL,
                          const double* const* AR, const double* const* C, const double* const* AC) {
  return guarded("imps_write_state_file", [&]() -> int {
    if (!path || nsites <= 0 || D <= 0 || d <= 0 || !AL || !AR || !C || !AC)
      return fail(IMPS_E_ARG, "imps_write_state_file: bad arguments");
    const size_t nA = (size_t)D * d * D, nC = (size_t)D * D;
    std::vector<std::pair<const double*, size_t>> blocks;
    for (int l = 0; l < nsites; ++l) {
      if (!AL[l] || !AR[l] || !C[l] || !AC[l]) return fail(IMPS_E_ARG, "imps_write_state_file: null tensor");
      blocks.push_back({AL[l], nA});
      blocks.push_back({AR[l], nA});
      blocks.push_back({C[l], nC});
      blocks.push_back({AC[l], nA});
    }
    return write_file(path, kStateMagic, nsites, D, d, blocks);
  });
}

int imps_state_file_info(const char* path, int* nsites, int* D, int* d) {
  return guarded("imps_state_file_info", [&]() -> int {
    return info_file(path, kStateMagic, nsites, D, d);
  });
}

int imps_read_state_file(const char* path, int nsites, int D, int d, double* const* AL, double* const* AR,
                         double* const* C, double* const* AC) {
  return guarded("imps_read_state_file", [&]() -> int {
    if (!path || nsites <= 0 || D <= 0 || d <= 0 || !AL || !AR || !C || !AC)
      return fail(IMPS_E_ARG, "imps_read_state_file: bad arguments");
    const size_t nA = (size_t)D * d * D, nC = (size_t)D * D;
    std::vector<std::pair<double*, size_t>> blocks;
    for (int l = 0; l < nsites; ++l) {
      if (!AL[l] || !AR[l] || !C[l] || !AC[l]) return fail(IMPS_E_ARG, "imps_read_state_file: null tensor");
      blocks.push_back({AL[l], nA});
      blocks.push_back({AR[l], nA});
      blocks.push_back({C[l], nC});
      blocks.push_back({AC[l], nA});
    }
    return read_file(path, kStateMagic, nsites, D, d, blocks);
  });
}

int imps_write_mpo_file(const char* path, int nsites, int w, int d, const double* const* W) {
  return guarded("imps_write_mpo_file", [&]() -> int {
    if (!path || nsites <= 0 || w <= 0 || d <= 0 || !W) return fail(IMPS_E_ARG, "imps_write_mpo_file: bad arguments");
    const size_t nW = (size_t)w * d * w * d;
    std::vector<std::pair<const double*, size_t>> blocks;
    for (int l = 0; l < nsites; ++l) {
      if (!W[l]) return fail(IMPS_E_ARG, "imps_write_mpo_file: null tensor");
      blocks.push_back({W[l], nW});
    }
    return write_file(path, kMpoMagic, nsites, w, d, blocks);
  });
}

int imps_mpo_file_info(const char* path, int* nsites, int* w, int* d) {
  return guarded("imps_mpo_file_info", [&]() -> int {
    return info_file(path, kMpoMagic, nsites, w, d);
  });
}

int imps_read_mpo_file(const char* path, int nsites, int w, int d, double* const* W) {
  return guarded("imps_read_mpo_file", [&]() -> int {
    if (!path || nsites <= 0 || w <= 0 || d <= 0 || !W) return fail(IMPS_E_ARG, "imps_read_mpo_file: bad arguments");
    const size_t nW = (size_t)w * d * w * d;
    std::vector<std::pair<double*, size_t>> blocks;
    for (int l = 0; l < nsites; ++l) {
      if (!W[l]) return fail(IMPS_E_ARG, "imps_read_mpo_file: null tensor");
      blocks.push_back({W[l], nW});
    }
    return read_file(path, kMpoMagic, nsites, w, d, blocks);
  });
}

int imps_checkpoint_state(imps_ctx* ctx, const char* path) {
  return guarded("imps_checkpoint_state", [&]() -> int {
    if (!ctx || !path) return IMPS_E_ARG;
    int N = 0, D = 0, d = 0;
    int st = imps_state_dims(ctx, &N, &D, &d);
    if (st != IMPS_OK) return st;
    const size_t nA = (size_t)D * d * D, nC = (size_t)D * D;
    std::vector<std::vector<double>> AL(N, std::vector<double>(nA)), AR = AL, AC = AL;
    std::vector<std::vector<double>> C(N, std::vector<double>(nC));
    std::vector<double*> pl(N), pr(N), pc(N), pa(N);
    for (int l = 0; l < N; ++l) {
      pl[l] = AL[l].data();
      pr[l] = AR[l].data();
      pc[l] = C[l].data();
      pa[l] = AC[l].data();
    }
    st = imps_get_state(ctx, pl.data(), pr.data(), pc.data(), pa.data());
    if (st != IMPS_OK) return st;
    st = imps_write_state_file(path, N, D, d, pl.data(), pr.data(), pc.data(), pa.data());
    if (st != IMPS_OK) imps::set_context_error(ctx, imps_last_error(nullptr));
    return st;
  });
}

int imps_resume_state(imps_ctx* ctx, const char* path) {
  return guarded("imps_resume_state", [&]() -> int {
    if (!ctx || !path) return IMPS_E_ARG;
    int N = 0, D = 0, d = 0;
    int st = imps_state_file_info(path, &N, &D, &d);
    if (st != IMPS_OK) {
      imps::set_context_error(ctx, imps_last_error(nullptr));
      return st;
    }
    // the file must fit the context's model before any buffer is sized from it: imps_set_state
    // reads D*d*D doubles per tensor with the model's d
    int mN = 0, md = 0;
    if (!imps::context_model(ctx, &mN, &md)) {
      imps::set_context_error(ctx, "imps_resume_state: MPO not set");
      return IMPS_E_STATE;
    }
    if (N != mN || d != md) {
      imps::set_context_error(ctx, "imps_resume_state: the file's unit cell (" + std::to_string(N) +
                                       ") or physical dimension (" + std::to_string(d) +
                                       ") differs from the model's (" + std::to_string(mN) + ", " +
                                       std::to_string(md) + ")");
      return IMPS_E_ARG;
    }
    const size_t nA = (size_t)D * d * D, nC = (size_t)D * D;
    std::vector<std::vector<double>> AL(N, std::vector<double>(nA)), AR = AL, AC = AL;
    std::vector<std::vector<double>> C(N, std::vector<double>(nC));
    std::vector<double*> pl(N), pr(N), pc(N), pa(N);
    for (int l = 0; l < N; ++l) {
      pl[l] = AL[l].data();
      pr[l] = AR[l].data();
      pc[l] = C[l].data();
      pa[l] = AC[l].data();
    }
    st = imps_read_state_file(path, N, D, d, pl.data(), pr.data(), pc.data(), pa.data());
    if (st != IMPS_OK) {
      imps::set_context_error(ctx, imps_last_error(nullptr));
      return st;
    }
    return imps_set_state(ctx, N, D, pl.data(), pr.data(), pc.data(), pa.data());
  });
}

}  // extern "C"
