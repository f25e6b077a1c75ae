/*
 * ecsr.c -- epsilon-neighbour lists on disk (.ecsr), so that later runs on the
 * same data skip the Smith-Waterman pass.  No counterpart in the reference.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mnhn_host.h"

static const char kMagic[8] = {'M', 'N', 'H', 'N', 'E', 'C', 'S', 'R'};

typedef struct {
  char magic[8];
  uint32_t version;
  int32_t n_rows;
  int32_t eps_max;
  int32_t reserved;
  int64_t nnz;
  uint64_t data_hash;
} ecsr_header;

int mnhn_csr_save(const mnhn_csr *csr, uint64_t data_hash, const char *path) {
  if (!csr || !path || csr->n_rows < 0 || csr->nnz < 0 || !csr->row_ptr ||
      (csr->nnz > 0 && (!csr->col || !csr->dist)))
    return MNHN_ERR_ARG;
  FILE *f = fopen(path, "wb");
  if (!f) return MNHN_ERR_ARG;
  ecsr_header h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kMagic, 8);
  h.version = 1;
  h.n_rows = csr->n_rows;
  h.eps_max = csr->eps_max;
  h.nnz = csr->nnz;
  h.data_hash = data_hash;
  const size_t nr = (size_t)csr->n_rows + 1, nz = (size_t)csr->nnz;
  int ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(csr->row_ptr, sizeof(int64_t), nr, f) == nr &&
           fwrite(csr->col, sizeof(int32_t), nz, f) == nz && fwrite(csr->dist, sizeof(uint16_t), nz, f) == nz;
  if (fclose(f) != 0) ok = 0;
  return ok ? MNHN_OK : MNHN_ERR_ARG;
}

int mnhn_csr_load(const char *path, uint64_t *data_hash, mnhn_csr *out) {
  if (!path || !out) return MNHN_ERR_ARG;
  memset(out, 0, sizeof(*out));
  FILE *f = fopen(path, "rb");
  if (!f) return MNHN_ERR_ARG;
  ecsr_header h;
  int ok = fread(&h, sizeof(h), 1, f) == 1 && memcmp(h.magic, kMagic, 8) == 0 && h.version == 1 &&
           h.n_rows >= 0 && h.nnz >= 0;
  if (ok) {
    const size_t nr = (size_t)h.n_rows + 1, nz = (size_t)h.nnz;
    out->row_ptr = (int64_t *)malloc(nr * sizeof(int64_t));
    out->col = (int32_t *)malloc((nz ? nz : 1) * sizeof(int32_t));
    out->dist = (uint16_t *)malloc((nz ? nz : 1) * sizeof(uint16_t));
    ok = out->row_ptr && out->col && out->dist && fread(out->row_ptr, sizeof(int64_t), nr, f) == nr &&
         fread(out->col, sizeof(int32_t), nz, f) == nz && fread(out->dist, sizeof(uint16_t), nz, f) == nz;
    /* the row pointers must describe nnz ascending entries */
    for (size_t r = 0; ok && r < nr; ++r)
      if (out->row_ptr[r] < (r ? out->row_ptr[r - 1] : 0) || out->row_ptr[r] > h.nnz) ok = 0;
    if (ok && out->row_ptr[0] != 0) ok = 0;
    if (ok && out->row_ptr[nr - 1] != h.nnz) ok = 0;
    for (size_t k = 0; ok && k < nz; ++k)
      if (out->col[k] < 0 || out->col[k] >= h.n_rows) ok = 0;
  }
  fclose(f);
  if (!ok) {
    free(out->row_ptr);
    free(out->col);
    free(out->dist);
    memset(out, 0, sizeof(*out));
    return MNHN_ERR_ARG;
  }
  out->n_rows = h.n_rows;
  out->eps_max = h.eps_max;
  out->nnz = h.nnz;
  if (data_hash) *data_hash = h.data_hash;
  return MNHN_OK;
}
