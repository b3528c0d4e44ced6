// Field files in the reference's on-disk format, behind the C ABI.
//
// MPISAVEX / MPILOADX (src/mod_fft.f90:2304-2642) write and read two files
// per field:
//   FN.info  Fortran unformatted sequential records (4-byte little-endian
//            length markers around each record): NDIMR,NDIMTH,NDIMX /
//            NRCHOPDIM,NTCHOPDIM,NXCHOPDIM / SPACE,LN / ZLEN,ELL / MINC,MKLINK
//            (INTEGER int32, REAL(8) float64);
//   FN       raw complex128 coefficients: a run on N1 x N2 processes writes
//            one block per process, the process's local array in column-major
//            order, at element offset INDEX * ARRAYSIZE.  The local extents
//            follow DECOMPOSE (src/mod_fft.f90:1400-1431): the first
//            MOD(N, P) processes of a split axis get one element more.
//            FFF and PFF split NXCHOPDIM over SUBCOMM_1 and NTCHOPDIM over
//            SUBCOMM_2; PPP and PFP split NDIMR over SUBCOMM_1 and NDIMX
//            over SUBCOMM_2; INDEX = P1 + P2 N1 when NXCHOPDIM < NTCHOPDIM,
//            else P2 + P1 N2; ARRAYSIZE uses the rounded-up extents.
// This library writes one block (N1 = N2 = 1); it reads files of any N1 x N2.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mlegs_internal.h"

#define RET_IF(x)            \
  do {                       \
    const int rc_ = (x);     \
    if (rc_) return rc_;     \
  } while (0)

namespace {

bool write_record(std::FILE *f, const void *p, int32_t n) {
  return std::fwrite(&n, 4, 1, f) == 1 && (n == 0 || std::fwrite(p, 1, (size_t)n, f) == (size_t)n) &&
         std::fwrite(&n, 4, 1, f) == 1;
}

// one logical record; compilers split records over 2 GiB into sub-records
// whose leading marker is negative while more follow
bool read_record(std::FILE *f, std::vector<char> &out) {
  out.clear();
  for (;;) {
    int32_t head = 0, tail = 0;
    if (std::fread(&head, 4, 1, f) != 1) return false;
    const bool more = head < 0;
    const size_t n = (size_t)(head < 0 ? -(int64_t)head : head);
    const size_t at = out.size();
    out.resize(at + n);
    if (n && std::fread(out.data() + at, 1, n, f) != n) return false;
    if (std::fread(&tail, 4, 1, f) != 1 || (tail < 0 ? -(int64_t)tail : tail) != (int64_t)n) return false;
    if (!more) return true;
  }
}

struct Info {
  int32_t ndimr, ndimth, ndimx, nrchopdim, ntchopdim, nxchopdim, space;
  double ln, zlen, ell;
  int32_t minc, mklink;
};

// (local size, start) of DECOMPOSE for process `rank` of `np`
void decompose(int n, int np, int rank, int &size, int &start) {
  const int q = n / np, r = n % np;
  if (r > rank) {
    size = q + 1;
    start = (q + 1) * rank;
  } else {
    size = q;
    start = q * rank + r;
  }
}

// full extents of a space and the axes split over SUBCOMM_1 and SUBCOMM_2
void split_axes(const Info &in, int space, int ext[3], int &ax1, int &ax2) {
  if (space == MLG_FFF || space == MLG_PFF) {
    ext[0] = space == MLG_FFF ? in.nrchopdim : in.ndimr;
    ext[1] = in.ntchopdim;
    ext[2] = in.nxchopdim;
    ax1 = 2;
    ax2 = 1;
  } else {
    ext[0] = in.ndimr;
    ext[1] = in.ndimth;
    ext[2] = in.ndimx;
    ax1 = 0;
    ax2 = 2;
  }
}

Info grid_info(const mlg_field *f, int space, double ln) {
  const mlg::Grid &G = f->ctx->g;
  Info in;
  in.ndimr = G.ndimr;
  in.ndimth = G.ndimth;
  in.ndimx = G.ndimx;
  in.nrchopdim = G.nrchopdim;
  in.ntchopdim = G.ntchopdim;
  in.nxchopdim = G.nxchopdim;
  in.space = space;
  in.ln = ln;
  in.zlen = G.p.zlen;
  in.ell = G.p.ell;
  in.minc = G.p.minc ? G.p.minc : 1;
  in.mklink = G.p.mklink;
  return in;
}

long space_elems(const Info &in, int space) {
  int ext[3], a1, a2;
  split_axes(in, space, ext, a1, a2);
  return (long)ext[0] * ext[1] * ext[2];
}

}  // namespace

// MPISAVEX of one process: FN.info and FN.
extern "C" int mlg_field_save(const mlg_field *f, const char *fn) {
  if (!f || !fn) return MLG_E_ARG;
  int space = 0;
  double ln = 0.0;
  RET_IF(mlg_field_download(f, nullptr, &space, nullptr));
  Info in = grid_info(f, space, 0.0);
  std::vector<double> e((size_t)2 * space_elems(in, space));
  RET_IF(mlg_field_download(f, e.data(), &space, &ln));
  in.ln = ln;
  const std::string info = std::string(fn) + ".info";
  std::FILE *fi = std::fopen(info.c_str(), "wb");
  if (!fi) return MLG_E_IO;
  const int32_t r0[3] = {in.ndimr, in.ndimth, in.ndimx}, r1[3] = {in.nrchopdim, in.ntchopdim, in.nxchopdim};
  char r2[12];
  std::memcpy(r2, &in.space, 4);
  std::memcpy(r2 + 4, &in.ln, 8);
  const double r3[2] = {in.zlen, in.ell};
  const int32_t r4[2] = {in.minc, in.mklink};
  bool ok = write_record(fi, r0, 12) && write_record(fi, r1, 12) && write_record(fi, r2, 12) && write_record(fi, r3, 16) &&
            write_record(fi, r4, 8);
  ok = (std::fclose(fi) == 0) && ok;
  if (!ok) return MLG_E_IO;
  std::FILE *fr = std::fopen(fn, "wb");
  if (!fr) return MLG_E_IO;
  ok = std::fwrite(e.data(), sizeof(double), e.size(), fr) == e.size();
  ok = (std::fclose(fr) == 0) && ok;
  return ok ? 0 : MLG_E_IO;
}

// MPILOADX: FN.info + FN written by a run on n1 x n2 processes.  A grid
// that differs from the file's returns 32 (DIMMISMATCH; the reference
// prints a warning and continues, but a field of another shape cannot be
// filled).
extern "C" int mlg_field_load(mlg_field *f, const char *fn, int n1, int n2) {
  if (!f || !fn || n1 < 1 || n2 < 1) return MLG_E_ARG;
  const std::string info = std::string(fn) + ".info";
  std::FILE *fi = std::fopen(info.c_str(), "rb");
  if (!fi) return MLG_E_IO;
  std::vector<char> r[5];
  bool ok = true;
  for (int i = 0; i < 5 && ok; i++) ok = read_record(fi, r[i]);
  std::fclose(fi);
  if (!ok || r[0].size() != 12 || r[1].size() != 12 || r[2].size() != 12 || r[3].size() != 16 || r[4].size() != 8)
    return MLG_E_IO;
  Info in;
  std::memcpy(&in.ndimr, r[0].data(), 4);
  std::memcpy(&in.ndimth, r[0].data() + 4, 4);
  std::memcpy(&in.ndimx, r[0].data() + 8, 4);
  std::memcpy(&in.nrchopdim, r[1].data(), 4);
  std::memcpy(&in.ntchopdim, r[1].data() + 4, 4);
  std::memcpy(&in.nxchopdim, r[1].data() + 8, 4);
  std::memcpy(&in.space, r[2].data(), 4);
  std::memcpy(&in.ln, r[2].data() + 4, 8);
  std::memcpy(&in.zlen, r[3].data(), 8);
  std::memcpy(&in.ell, r[3].data() + 8, 8);
  std::memcpy(&in.minc, r[4].data(), 4);
  std::memcpy(&in.mklink, r[4].data() + 4, 4);
  if (in.space < 0 || in.space > 3) return MLG_E_IO;
  const Info mine = grid_info(f, in.space, in.ln);
  if (in.ndimr != mine.ndimr || in.ndimth != mine.ndimth || in.ndimx != mine.ndimx || in.nrchopdim != mine.nrchopdim ||
      in.ntchopdim != mine.ntchopdim || in.nxchopdim != mine.nxchopdim)
    return 32;
  int ext[3], ax1, ax2;
  split_axes(in, in.space, ext, ax1, ax2);
  long arraysize = 1;
  for (int a = 0; a < 3; a++) arraysize *= a == ax1 ? (ext[a] + n1 - 1) / n1 : a == ax2 ? (ext[a] + n2 - 1) / n2 : ext[a];
  std::vector<double> e((size_t)2 * ext[0] * ext[1] * ext[2], 0.0);
  std::FILE *fr = std::fopen(fn, "rb");
  if (!fr) return MLG_E_IO;
  std::vector<double> blk;
  for (int p1 = 0; p1 < n1 && ok; p1++)
    for (int p2 = 0; p2 < n2 && ok; p2++) {
      const long index = in.nxchopdim < in.ntchopdim ? p1 + (long)p2 * n1 : p2 + (long)p1 * n2;
      int sz[3], st[3];
      for (int a = 0; a < 3; a++) {
        if (a == ax1) decompose(ext[a], n1, p1, sz[a], st[a]);
        else if (a == ax2) decompose(ext[a], n2, p2, sz[a], st[a]);
        else {
          sz[a] = ext[a];
          st[a] = 0;
        }
      }
      const size_t nloc = (size_t)sz[0] * sz[1] * sz[2];
      blk.resize(2 * nloc);
      ok = std::fseek(fr, (long)(index * arraysize * 16), SEEK_SET) == 0 &&
           std::fread(blk.data(), 16, nloc, fr) == nloc;
      for (int k = 0; ok && k < sz[2]; k++)
        for (int j = 0; j < sz[1]; j++) {
          const size_t src = 2 * ((size_t)sz[0] * (j + (size_t)sz[1] * k));
          const size_t dst = 2 * ((size_t)st[0] + (size_t)ext[0] * ((st[1] + j) + (size_t)ext[1] * (st[2] + k)));
          std::memcpy(&e[dst], &blk[src], sizeof(double) * 2 * sz[0]);
        }
    }
  std::fclose(fr);
  if (!ok) return MLG_E_IO;
  return mlg_field_upload(f, e.data(), in.space, in.ln);
}

// The m = 0 columns E(:, 1, k), k < nk, of a field in its current space
// (INSPECT, src/mod_march.f90:3743-3775, prints them): out[2 (r + nrad k)],
// *nrad = the space's first extent.  Synchronising.
extern "C" int mlg_field_m0_columns(const mlg_field *f, int nk, double *out, int *nrad) {
  if (!f || nk < 1 || !nrad) return MLG_E_ARG;
  const mlg::Grid &G = f->ctx->g;
  const int space = f->space;
  const int n1 = space == MLG_FFF ? G.nrchopdim : G.ndimr;
  const int n2 = (space == MLG_FFF || space == MLG_PFF) ? G.ntchopdim : G.ndimth;
  const int n3 = (space == MLG_FFF || space == MLG_PFF) ? G.nxchopdim : G.ndimx;
  *nrad = n1;
  if (!out) return 0;
  if (nk > n3) return MLG_E_ARG;
  MLG_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
  for (int k = 0; k < nk; k++)
    MLG_CHECK_HIP(hipMemcpy(out + (size_t)2 * n1 * k, f->e + (size_t)n1 * n2 * k, sizeof(double2) * (size_t)n1,
                            hipMemcpyDeviceToHost));
  return 0;
}
