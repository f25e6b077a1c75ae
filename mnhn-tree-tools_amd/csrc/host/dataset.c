/*
 * dataset.c -- FASTA reader with the reference's parsing rules
 * (dataset_from_fasta, dataset.c:264-344 of the reference):
 *   - a line whose first byte is '>' opens a new record (its text is ignored);
 *   - any other line contributes its first whitespace-delimited token,
 *     upper-cased (toupper) and appended to the current record;
 *   - a line without a token appends the previous token again, because the
 *     reference's sscanf("%s") leaves its line buffer unchanged;
 *   - every sequence is NUL-terminated; non-ACGT bytes are kept as they are.
 * Unlike the reference, tokens are not limited to 1,999 bytes, and text
 * before the first header is ignored (the reference writes it out of bounds).
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <unistd.h>
#include <string.h>

#include "mnhn_host.h"

static void *xrealloc(void *p, size_t n) {
  void *q = realloc(p, n ? n : 1);
  if (!q) {
    printf("out of memory reading FASTA\n");
    exit(1);
  }
  return q;
}

dataset dataset_from_fasta(FILE *in) {
  dataset ds;
  memset(&ds, 0, sizeof(ds));
  char *line = NULL;
  size_t cap = 0;
  ssize_t got;
  int records = 0;

  rewind(in);
  while ((got = getline(&line, &cap, in)) != -1)
    if (line[0] == '>') records++;
  rewind(in);

  ds.n_values = records;
  ds.sequences = (char **)xrealloc(NULL, sizeof(char *) * (size_t)records);
  ds.sequence_lengths = (size_t *)xrealloc(NULL, sizeof(size_t) * (size_t)records);
  size_t *caps = (size_t *)xrealloc(NULL, sizeof(size_t) * (size_t)records);

  char *tok = NULL; /* last token seen (sscanf's line buffer) */
  size_t tok_len = 0, tok_cap = 0;
  int cur = -1;
  while ((got = getline(&line, &cap, in)) != -1) {
    if (line[0] == '>') {
      cur++;
      caps[cur] = 1024;
      ds.sequences[cur] = (char *)xrealloc(NULL, caps[cur]);
      ds.sequences[cur][0] = 0;
      ds.sequence_lengths[cur] = 0;
      continue;
    }
    /* first token of the line, as sscanf("%s") would read it */
    size_t k = 0;
    while (k < (size_t)got && line[k] != 0 && isspace((unsigned char)line[k])) k++;
    if (k < (size_t)got && line[k] != 0) {
      size_t e = k;
      while (e < (size_t)got && line[e] != 0 && !isspace((unsigned char)line[e])) e++;
      tok_len = e - k;
      if (tok_len + 1 > tok_cap) {
        tok_cap = 2 * (tok_len + 1);
        tok = (char *)xrealloc(tok, tok_cap);
      }
      for (size_t t = 0; t < tok_len; t++) tok[t] = (char)toupper((unsigned char)line[k + t]);
    }
    if (cur < 0 || tok_len == 0) continue;
    size_t L = ds.sequence_lengths[cur];
    if (L + tok_len + 1 > caps[cur]) {
      while (L + tok_len + 1 > caps[cur]) caps[cur] *= 2;
      ds.sequences[cur] = (char *)xrealloc(ds.sequences[cur], caps[cur]);
    }
    memcpy(ds.sequences[cur] + L, tok, tok_len);
    ds.sequences[cur][L + tok_len] = 0;
    ds.sequence_lengths[cur] = L + tok_len;
  }
  for (int i = 0; i < records; i++) {
    ds.sequences[i] = (char *)xrealloc(ds.sequences[i], ds.sequence_lengths[i] + 1);
    if (ds.max_sequence_length < ds.sequence_lengths[i])
      ds.max_sequence_length = ds.sequence_lengths[i];
  }
  free(caps);
  free(tok);
  free(line);
  return ds;
}

void free_sequences_from_dataset(dataset ds) {
  for (int i = 0; i < ds.n_values; i++) free(ds.sequences[i]);
  free(ds.sequence_lengths);
  free(ds.sequences);
}

/* dataset.c:427-469: n_values rows of `dimensions` whitespace-separated
 * floats; values[dimension][row].  A short row prints the reference's message
 * and exits. */
void load_projections_from_file_into_dataset(FILE *projections, size_t dimensions, dataset *ds) {
  ds->n_dimensions = (int)dimensions;
  ds->values = (float **)malloc(sizeof(float *) * (dimensions ? dimensions : 1));
  for (size_t i = 0; i < dimensions; i++) {
    ds->values[i] = (float *)malloc(sizeof(float) * (size_t)(ds->n_values + 8));
    if (!ds->values[i]) {
      printf("Could not allocate memory to load projections \n");
      _exit(1);
    }
  }
  for (int i = 0; i < ds->n_values; i++) {
    size_t counter = 0;
    for (size_t j = 0; j < dimensions; j++) counter += (size_t)fscanf(projections, "%f", ds->values[j] + i);
    if (counter != dimensions) {
      printf("Error reading projections file \n");
      _exit(1);
    }
  }
}

/* dataset.c:629-635 */
void free_values_from_dataset(dataset ds) {
  for (int i = 0; i < ds.n_dimensions; i++) free(ds.values[i]);
  free(ds.values);
}

/* dataset.c:646-649 */
void free_dataset(dataset ds) {
  free_values_from_dataset(ds);
  free_sequences_from_dataset(ds);
}

/* dataset.c:346-375: tabs before the first line break are the features; the
 * line breaks from that one to the end of the file are the samples.  Read in
 * blocks instead of whole; a file without a line break has 0 samples (the
 * reference reads past its buffer there). */
data_shape shape_from_kmer_file(int infile) {
  data_shape s = {0, 0};
  char buf[1 << 16];
  ssize_t got;
  int first_line = 1;
  if (lseek(infile, 0, SEEK_SET) < 0) return s;
  while ((got = read(infile, buf, sizeof(buf))) > 0) {
    for (ssize_t i = 0; i < got; i++) {
      if (first_line) {
        if (buf[i] == '\t') s.n_features++;
        if (buf[i] == '\n') {
          first_line = 0;
          s.n_samples++;
        }
      } else if (buf[i] == '\n') {
        s.n_samples++;
      }
    }
  }
  return s;
}

/* dataset.c:378-425: per sample one whitespace-delimited label token, then
 * n_features values as fscanf("%f") reads them (strtof accepts the same
 * decimal, hexadecimal, inf and nan forms and rounds the same way).  The file
 * is parsed from memory; values missing at the end of the file stay 0. */
dataset load_kmer_from_file_into_dataset(FILE *in_file, data_shape shape) {
  dataset ds;
  memset(&ds, 0, sizeof(ds));
  ds.n_dimensions = (int)shape.n_features;
  ds.n_values = (int)shape.n_samples;
  ds.values = (float **)malloc(sizeof(float *) * (shape.n_features ? shape.n_features : 1));
  for (size_t k = 0; k < shape.n_features; k++) {
    ds.values[k] = (float *)calloc((size_t)ds.n_values + 8, sizeof(float));
    if (!ds.values[k]) {
      printf("Error allocating aligned memory for kmers \n");
      _exit(1);
    }
  }
  rewind(in_file);
  size_t cap = 1 << 20, len = 0;
  char *text = (char *)xrealloc(NULL, cap);
  size_t r;
  while ((r = fread(text + len, 1, cap - len - 1, in_file)) > 0) {
    len += r;
    if (cap - len - 1 == 0) {
      cap *= 2;
      text = (char *)xrealloc(text, cap);
    }
  }
  text[len] = 0;
  char *p = text;
  for (size_t i = 0; i < shape.n_samples; i++) {
    while (*p && isspace((unsigned char)*p)) p++; /* label token */
    while (*p && !isspace((unsigned char)*p)) p++;
    for (size_t k = 0; k < shape.n_features; k++) {
      char *end;
      const float v = strtof(p, &end);
      if (end == p) goto done;
      ds.values[k][i] = v;
      p = end;
    }
  }
done:
  free(text);
  return ds;
}

