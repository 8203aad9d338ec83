// h3dmg_tecbin.cpp -- Hydro3D's binary restart file of one block on the host.
//
// Reference: the writer `tecbin`, src_v70/post.for:324-375, and the restart
// reader, src_v70/initial.for:715-747.  Both use Fortran unformatted
// sequential access: every WRITE is one record, stored as a 4-byte payload
// length, the payload, and the same length again (the convention of the
// Fortran runtimes the reference is built with; records here are far below
// the 2 GiB at which runtimes split a record).  The records are
//   toti, totj, totk                  three default INTEGERs (4 bytes each)
//   pl                                one INTEGER
//   inind, jnind, knind               three INTEGERs
//   per cell, k = 1..totk, j = 1..totj, i = 1..toti (i fastest):
//     x(i), y(j), z(k) and the 20 fields at (i, j, k)       23 doubles
// with the fields in the order of h3dmg.h's H3DMG_TECBIN_* indices.
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/h3dmg.h"

namespace {

thread_local std::string g_tecbin_error;

h3dmg_status tfail(h3dmg_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
h3dmg_status tfail(h3dmg_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_tecbin_error = buf;
  return st;
}

constexpr int NCOL = 3 + H3DMG_TECBIN_FIELDS;  // 23 doubles per cell record

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// one record: length, payload, length
bool put_record(FILE *f, const void *data, uint32_t bytes) {
  return fwrite(&bytes, 4, 1, f) == 1 && fwrite(data, 1, bytes, f) == bytes &&
         fwrite(&bytes, 4, 1, f) == 1;
}

// one record of exactly `bytes` payload bytes; false with the error set
// otherwise (end of file, or a record of another length)
bool get_record(FILE *f, void *data, uint32_t bytes, const char *what, h3dmg_status *st) {
  uint32_t head = 0, tail = 0;
  if (fread(&head, 4, 1, f) != 1) {
    *st = tfail(H3DMG_ERR_ARG, "tecbin: end of file before the %s record", what);
    return false;
  }
  if (head != bytes) {
    *st = tfail(H3DMG_ERR_ARG, "tecbin: the %s record holds %u bytes, expected %u", what, head,
                bytes);
    return false;
  }
  if (fread(data, 1, bytes, f) != bytes || fread(&tail, 4, 1, f) != 1) {
    *st = tfail(H3DMG_ERR_ARG, "tecbin: the %s record is truncated", what);
    return false;
  }
  if (tail != head) {
    *st = tfail(H3DMG_ERR_ARG, "tecbin: the %s record's length markers differ (%u, %u)", what,
                head, tail);
    return false;
  }
  return true;
}

h3dmg_status read_header(FILE *f, h3dmg_tecbin_header *hd) {
  h3dmg_status st = H3DMG_OK;
  int32_t tt[3], pl, ind[3];
  if (!get_record(f, tt, sizeof(tt), "toti,totj,totk", &st)) return st;
  if (!get_record(f, &pl, sizeof(pl), "pl", &st)) return st;
  if (!get_record(f, ind, sizeof(ind), "inind,jnind,knind", &st)) return st;
  for (int a = 0; a < 3; ++a)
    if (tt[a] <= 0) return tfail(H3DMG_ERR_ARG, "tecbin: dimension %d is %d", a + 1, tt[a]);
  for (int a = 0; a < 3; ++a) {
    hd->tt[a] = tt[a];
    hd->ind[a] = ind[a];
  }
  hd->pl = pl;
  return H3DMG_OK;
}

}  // namespace

extern "C" {

const char *h3dmg_tecbin_error(void) { return g_tecbin_error.c_str(); }

h3dmg_status h3dmg_tecbin_write(const char *path, const h3dmg_tecbin_header *hd, const double *x,
                                const double *y, const double *z, const double *const *fields) {
  if (!path || !hd || !x || !y || !z || !fields)
    return tfail(H3DMG_ERR_ARG, "tecbin: null argument");
  for (int q = 0; q < H3DMG_TECBIN_FIELDS; ++q)
    if (!fields[q]) return tfail(H3DMG_ERR_ARG, "tecbin: field %d is null", q);
  for (int a = 0; a < 3; ++a)
    if (hd->tt[a] <= 0) return tfail(H3DMG_ERR_ARG, "tecbin: dimension %d is %d", a + 1, hd->tt[a]);
  File F;
  F.f = fopen(path, "wb");
  if (!F.f) return tfail(H3DMG_ERR_ARG, "tecbin: cannot open %s for writing", path);
  std::vector<char> iobuf(1 << 20);
  setvbuf(F.f, iobuf.data(), _IOFBF, iobuf.size());
  const int32_t tt[3] = {hd->tt[0], hd->tt[1], hd->tt[2]};
  const int32_t pl = hd->pl;
  const int32_t ind[3] = {hd->ind[0], hd->ind[1], hd->ind[2]};
  bool ok = put_record(F.f, tt, sizeof(tt)) && put_record(F.f, &pl, sizeof(pl)) &&
            put_record(F.f, ind, sizeof(ind));
  const size_t ni = hd->tt[0], nj = hd->tt[1], nk = hd->tt[2];
  double rec[NCOL];
  for (size_t k = 0; ok && k < nk; ++k)
    for (size_t j = 0; ok && j < nj; ++j)
      for (size_t i = 0; ok && i < ni; ++i) {
        const size_t c = (k * nj + j) * ni + i;
        rec[0] = x[i];
        rec[1] = y[j];
        rec[2] = z[k];
        for (int q = 0; q < H3DMG_TECBIN_FIELDS; ++q) rec[3 + q] = fields[q][c];
        ok = put_record(F.f, rec, sizeof(rec));
      }
  ok = ok && fflush(F.f) == 0;
  if (!ok) return tfail(H3DMG_ERR_ARG, "tecbin: write to %s failed", path);
  return H3DMG_OK;
}

h3dmg_status h3dmg_tecbin_read_header(const char *path, h3dmg_tecbin_header *hd) {
  if (!path || !hd) return tfail(H3DMG_ERR_ARG, "tecbin: null argument");
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) return tfail(H3DMG_ERR_ARG, "tecbin: cannot open %s", path);
  return read_header(F.f, hd);
}

h3dmg_status h3dmg_tecbin_read(const char *path, int pl, double *x, double *y, double *z,
                               double *const *fields) {
  if (!path) return tfail(H3DMG_ERR_ARG, "tecbin: null path");
  File F;
  F.f = fopen(path, "rb");
  if (!F.f) return tfail(H3DMG_ERR_ARG, "tecbin: cannot open %s", path);
  std::vector<char> iobuf(1 << 20);
  setvbuf(F.f, iobuf.data(), _IOFBF, iobuf.size());
  h3dmg_tecbin_header hd{};
  h3dmg_status st = read_header(F.f, &hd);
  if (st) return st;
  // initial.for:723-728: a restart with another number of ghost layers stops
  if (pl >= 0 && hd.pl != pl)
    return tfail(H3DMG_ERR_ARG, "tecbin: different number of overlapping layers (%d, file %d)",
                 pl, hd.pl);
  const size_t ni = hd.tt[0], nj = hd.tt[1], nk = hd.tt[2];
  double rec[NCOL];
  for (size_t k = 0; k < nk; ++k)
    for (size_t j = 0; j < nj; ++j)
      for (size_t i = 0; i < ni; ++i) {
        if (!get_record(F.f, rec, sizeof(rec), "cell", &st)) return st;
        const size_t c = (k * nj + j) * ni + i;
        if (x && j == 0 && k == 0) x[i] = rec[0];
        if (y && i == 0 && k == 0) y[j] = rec[1];
        if (z && i == 0 && j == 0) z[k] = rec[2];
        if (fields)
          for (int q = 0; q < H3DMG_TECBIN_FIELDS; ++q)
            if (fields[q]) fields[q][c] = rec[3 + q];
      }
  return H3DMG_OK;
}

}  // extern "C"
