/*
 * A small HDF5 writer and reader for the driver's files (see emc_io.h).
 *
 * File layout written (all addresses absolute, 8-byte offsets and lengths,
 * little endian, every structure 8-byte aligned):
 *   superblock version 0 with the root group's symbol-table entry;
 *   root object header (version 1) holding one symbol-table message;
 *   v1 B-tree (group node, level 0) over symbol-table nodes of up to 8
 *     entries each, names in increasing strcmp order;
 *   local heap with the names;
 *   per dataset: a version-1 object header (dataspace v1, datatype, fill
 *     value v2, layout v3 contiguous) followed by its raw data.
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "emc_io.h"

#define UNDEF_ADDR 0xffffffffffffffffULL
enum { LEAF_K = 4, NODE_K = 16, MAX_SETS = 64 };

static char g_err[256];

static int fail(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return -1;
}

const char *emc_io_error(void) { return g_err; }

/* ---- growable little-endian byte buffer ---- */
typedef struct {
  unsigned char *p;
  size_t n, cap;
  int oom;
} buf_t;

static void reserve(buf_t *b, size_t more) {
  if (b->oom || b->n + more <= b->cap) return;
  size_t c = b->cap ? b->cap : 4096;
  while (c < b->n + more) c *= 2;
  unsigned char *q = realloc(b->p, c);
  if (!q) {
    b->oom = 1;
    return;
  }
  b->p = q;
  b->cap = c;
}
static void put(buf_t *b, const void *src, size_t n) {
  reserve(b, n);
  if (b->oom) return;
  memcpy(b->p + b->n, src, n);
  b->n += n;
}
static void put_le(buf_t *b, uint64_t v, int bytes) { /* bytes <= 8 */
  unsigned char t[8];
  for (int i = 0; i < bytes; i++) t[i] = (unsigned char)(v >> (8 * i));
  put(b, t, bytes);
}
static void pad_to(buf_t *b, size_t n) {
  static const unsigned char z[64] = {0};
  while (b->n < n) put(b, z, n - b->n < sizeof z ? n - b->n : sizeof z);
}
static void align8(buf_t *b) { pad_to(b, (b->n + 7) & ~(size_t)7); }
static void set_le(buf_t *b, size_t at, uint64_t v, int bytes) {
  if (b->oom) return;
  for (int i = 0; i < bytes; i++) b->p[at + i] = (unsigned char)(v >> (8 * i));
}

static int type_size(int type) { return type == EMC_H5_F64 ? 8 : 4; }

static long long n_elems(const emc_h5_dataset *d) {
  long long n = 1;
  for (int k = 0; k < d->rank; k++) n *= d->dims[k];
  return n;
}

/* one object-header message (version-1 header): type, size, flags, data
 * padded to 8 bytes */
static void msg_begin(buf_t *b, int type, size_t *size_at) {
  put_le(b, (uint64_t)type, 2);
  *size_at = b->n;
  put_le(b, 0, 2);
  put_le(b, 0, 1);
  put_le(b, 0, 3);
}
static void msg_end(buf_t *b, size_t size_at) {
  align8(b);
  set_le(b, size_at, b->n - (size_at + 6), 2);
}

static void datatype_msg(buf_t *b, int type) {
  size_t at;
  msg_begin(b, 0x0003, &at);
  if (type == EMC_H5_I32) {
    put_le(b, 0x10, 1);           /* version 1, class 0 (fixed point) */
    put_le(b, 0x08, 1);           /* little endian, signed */
    put_le(b, 0, 2);
    put_le(b, 4, 4);              /* size */
    put_le(b, 0, 2);              /* bit offset */
    put_le(b, 32, 2);             /* bit precision */
  } else {
    const int dbl = type == EMC_H5_F64;
    put_le(b, 0x11, 1);           /* version 1, class 1 (floating point) */
    put_le(b, 0x20, 1);           /* little endian, mantissa normalised (implied msb) */
    put_le(b, dbl ? 63 : 31, 1);  /* sign bit position */
    put_le(b, 0, 1);
    put_le(b, dbl ? 8 : 4, 4);    /* size */
    put_le(b, 0, 2);              /* bit offset */
    put_le(b, dbl ? 64 : 32, 2);  /* bit precision */
    put_le(b, dbl ? 52 : 23, 1);  /* exponent location */
    put_le(b, dbl ? 11 : 8, 1);   /* exponent size */
    put_le(b, 0, 1);              /* mantissa location */
    put_le(b, dbl ? 52 : 23, 1);  /* mantissa size */
    put_le(b, dbl ? 1023 : 127, 4); /* exponent bias */
  }
  msg_end(b, at);
}

/* Dataset object header + raw data at b->n (8-aligned); returns the header
 * address. */
static uint64_t write_dataset(buf_t *b, const emc_h5_dataset *d) {
  align8(b);
  const uint64_t oh = b->n;
  const long long n = n_elems(d);
  const uint64_t bytes = (uint64_t)n * type_size(d->type);
  put_le(b, 1, 1); /* object header version 1 */
  put_le(b, 0, 1);
  put_le(b, 4, 2); /* messages */
  put_le(b, 1, 4); /* reference count */
  const size_t hsize_at = b->n;
  put_le(b, 0, 4); /* header data size, set below */
  put_le(b, 0, 4); /* alignment */
  const size_t h0 = b->n;
  size_t at;
  /* dataspace, version 1 (a scalar as rank 0) */
  msg_begin(b, 0x0001, &at);
  put_le(b, 1, 1);
  put_le(b, (uint64_t)d->rank, 1);
  put_le(b, 0, 1); /* no maximum dims */
  put_le(b, 0, 1);
  put_le(b, 0, 4);
  for (int k = 0; k < d->rank; k++) put_le(b, (uint64_t)d->dims[k], 8);
  msg_end(b, at);
  datatype_msg(b, d->type);
  /* fill value, version 2: allocation early, written if set, undefined */
  msg_begin(b, 0x0005, &at);
  put_le(b, 2, 1);
  put_le(b, 1, 1);
  put_le(b, 2, 1);
  put_le(b, 0, 1);
  msg_end(b, at);
  /* layout, version 3, contiguous: address and size */
  msg_begin(b, 0x0008, &at);
  put_le(b, 3, 1);
  put_le(b, 1, 1);
  const size_t addr_at = b->n;
  put_le(b, 0, 8);
  put_le(b, bytes, 8);
  msg_end(b, at);
  set_le(b, hsize_at, b->n - h0, 4);
  align8(b);
  set_le(b, addr_at, b->n, 8);
  put(b, d->data, (size_t)bytes);
  align8(b);
  return oh;
}

static const emc_h5_dataset *g_sort_sets;
static int cmp_idx(const void *a, const void *b) {
  return strcmp(g_sort_sets[*(const int *)a].name, g_sort_sets[*(const int *)b].name);
}

int emc_h5_write(const char *path, const emc_h5_dataset *sets, int n) {
  if (!path || n < 0 || n > MAX_SETS || (n > 0 && !sets)) return fail("emc_h5_write: invalid arguments");
  for (int i = 0; i < n; i++) {
    const emc_h5_dataset *d = &sets[i];
    if (!d->name || !d->name[0] || strchr(d->name, '/') || d->rank < 0 || d->rank > 3 || !d->data ||
        (d->type != EMC_H5_F32 && d->type != EMC_H5_I32 && d->type != EMC_H5_F64))
      return fail("emc_h5_write: dataset %d invalid", i);
    for (int k = 0; k < d->rank; k++)
      if (d->dims[k] < 0) return fail("emc_h5_write: dataset %s has a negative dimension", d->name);
    for (int j = 0; j < i; j++)
      if (!strcmp(sets[j].name, d->name)) return fail("emc_h5_write: dataset %s given twice", d->name);
  }
  int order[MAX_SETS];
  for (int i = 0; i < n; i++) order[i] = i;
  g_sort_sets = sets;
  qsort(order, (size_t)n, sizeof(int), cmp_idx);
  const int nsnod = n > 0 ? (n + 2 * LEAF_K - 1) / (2 * LEAF_K) : 1;
  if (nsnod > 2 * NODE_K) return fail("emc_h5_write: too many datasets");

  buf_t b = {0};
  /* superblock v0 */
  put(&b, "\211HDF\r\n\032\n", 8);
  put_le(&b, 0, 1); /* superblock version */
  put_le(&b, 0, 1); /* free-space version */
  put_le(&b, 0, 1); /* root group symbol table version */
  put_le(&b, 0, 1);
  put_le(&b, 0, 1); /* shared header message version */
  put_le(&b, 8, 1); /* size of offsets */
  put_le(&b, 8, 1); /* size of lengths */
  put_le(&b, 0, 1);
  put_le(&b, LEAF_K, 2);
  put_le(&b, NODE_K, 2);
  put_le(&b, 0, 4);          /* consistency flags */
  put_le(&b, 0, 8);          /* base address */
  put_le(&b, UNDEF_ADDR, 8); /* free space */
  const size_t eof_at = b.n;
  put_le(&b, 0, 8);          /* end of file, set at the end */
  put_le(&b, UNDEF_ADDR, 8); /* driver information */
  /* root symbol table entry */
  put_le(&b, 0, 8);
  const size_t root_oh_at = b.n;
  put_le(&b, 0, 8);
  put_le(&b, 1, 4); /* cache type 1: scratch holds B-tree and heap addresses */
  put_le(&b, 0, 4);
  const size_t root_btree_at = b.n;
  put_le(&b, 0, 8);
  const size_t root_heap_at = b.n;
  put_le(&b, 0, 8);
  align8(&b);

  /* root object header: one symbol-table message */
  const uint64_t root_oh = b.n;
  put_le(&b, 1, 1);
  put_le(&b, 0, 1);
  put_le(&b, 1, 2);
  put_le(&b, 1, 4);
  put_le(&b, 24, 4);
  put_le(&b, 0, 4);
  size_t at;
  msg_begin(&b, 0x0011, &at);
  const size_t stab_btree_at = b.n;
  put_le(&b, 0, 8);
  const size_t stab_heap_at = b.n;
  put_le(&b, 0, 8);
  msg_end(&b, at);

  /* local heap: "" at 0, then the names, each padded to 8 bytes */
  size_t name_off[MAX_SETS];
  size_t heap_size = 8;
  for (int j = 0; j < n; j++) {
    name_off[j] = heap_size;
    heap_size += (strlen(sets[order[j]].name) + 1 + 7) & ~(size_t)7;
  }
  const uint64_t heap = b.n;
  put(&b, "HEAP", 4);
  put_le(&b, 0, 1);
  put_le(&b, 0, 3);
  put_le(&b, heap_size, 8);
  put_le(&b, UNDEF_ADDR, 8); /* no free block */
  put_le(&b, b.n + 8, 8);    /* data segment right after */
  const size_t heap_data = b.n;
  put_le(&b, 0, 8);
  for (int j = 0; j < n; j++) {
    put(&b, sets[order[j]].name, strlen(sets[order[j]].name) + 1);
    align8(&b);
  }
  pad_to(&b, heap_data + heap_size);

  /* B-tree node (full size: 2 NODE_K children) */
  const uint64_t btree = b.n;
  put(&b, "TREE", 4);
  put_le(&b, 0, 1); /* group node */
  put_le(&b, 0, 1); /* leaf level */
  put_le(&b, (uint64_t)nsnod, 2);
  put_le(&b, UNDEF_ADDR, 8);
  put_le(&b, UNDEF_ADDR, 8);
  const size_t keys_at = b.n;
  pad_to(&b, keys_at + (2 * NODE_K + 1) * 8 + 2 * NODE_K * 8);

  /* symbol-table nodes (full size: 2 LEAF_K entries), then the datasets */
  uint64_t snod[2 * NODE_K];
  size_t ent_at[MAX_SETS] = {0};
  for (int s = 0; s < nsnod; s++) {
    snod[s] = b.n;
    const int j0 = s * 2 * LEAF_K, j1 = j0 + 2 * LEAF_K < n ? j0 + 2 * LEAF_K : n;
    put(&b, "SNOD", 4);
    put_le(&b, 1, 1);
    put_le(&b, 0, 1);
    put_le(&b, (uint64_t)(j1 - j0), 2);
    const size_t e0 = b.n;
    for (int j = j0; j < j1; j++) {
      put_le(&b, name_off[j], 8);
      ent_at[j] = b.n;
      put_le(&b, 0, 8); /* object header address, set below */
      put_le(&b, 0, 4); /* cache type 0 */
      put_le(&b, 0, 4);
      put_le(&b, 0, 8); /* scratch pad, 16 bytes */
      put_le(&b, 0, 8);
    }
    pad_to(&b, e0 + 2 * LEAF_K * 40);
  }
  for (int j = 0; j < n; j++) {
    const uint64_t oh = write_dataset(&b, &sets[order[j]]);
    set_le(&b, ent_at[j], oh, 8);
  }
  /* keys: key 0 = "" (offset 0); key s+1 = the last name of node s */
  set_le(&b, keys_at, 0, 8);
  for (int s = 0; s < nsnod; s++) {
    const int last = (s + 1) * 2 * LEAF_K < n ? (s + 1) * 2 * LEAF_K - 1 : n - 1;
    set_le(&b, keys_at + 8 + (size_t)s * 16, snod[s], 8);
    set_le(&b, keys_at + 16 + (size_t)s * 16, n > 0 ? name_off[last] : 0, 8);
  }
  set_le(&b, root_oh_at, root_oh, 8);
  set_le(&b, root_btree_at, btree, 8);
  set_le(&b, root_heap_at, heap, 8);
  set_le(&b, stab_btree_at, btree, 8);
  set_le(&b, stab_heap_at, heap, 8);
  set_le(&b, eof_at, b.n, 8);
  if (b.oom) {
    free(b.p);
    return fail("emc_h5_write: out of memory");
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    free(b.p);
    return fail("emc_h5_write: cannot create %s", path);
  }
  const size_t w = fwrite(b.p, 1, b.n, f);
  const int closed = fclose(f);
  free(b.p);
  if (w != b.n || closed != 0) return fail("emc_h5_write: short write to %s", path);
  return 0;
}

/* ---- reader ---- */
typedef struct {
  const unsigned char *p;
  size_t n;
} file_t;

static int rd(const file_t *f, uint64_t at, size_t len, const unsigned char **out) {
  if (at > f->n || len > f->n - at) return -1;
  *out = f->p + at;
  return 0;
}
static uint64_t le(const unsigned char *q, int bytes) {
  uint64_t v = 0;
  for (int i = 0; i < bytes; i++) v |= (uint64_t)q[i] << (8 * i);
  return v;
}

typedef struct {
  int rank;
  long long dims[3];
  int cls, size, is_signed; /* datatype: class 0 integer / 1 float, bytes */
  uint64_t addr, bytes;     /* contiguous data */
  const unsigned char *compact;
  int have_space, have_type, have_layout;
} dset_t;

static int parse_messages(const file_t *f, uint64_t at, uint64_t len, dset_t *d, int depth) {
  if (depth > 8) return fail("object header continuation too deep");
  uint64_t pos = at, end = at + len;
  while (pos + 8 <= end) {
    const unsigned char *h;
    if (rd(f, pos, 8, &h)) return fail("object header out of the file");
    const int type = (int)le(h, 2), size = (int)le(h + 2, 2);
    const unsigned char *m;
    if (rd(f, pos + 8, (size_t)size, &m)) return fail("object header message out of the file");
    if (type == 0x0001) {
      const int ver = m[0];
      d->rank = m[1];
      if (d->rank > 3) return fail("dataset rank %d > 3", d->rank);
      const int off = ver == 1 ? 8 : 4;
      if (ver != 1 && ver != 2) return fail("dataspace version %d", ver);
      if (ver == 2 && m[3] == 2) d->rank = 0; /* null dataspace */
      for (int k = 0; k < d->rank; k++) d->dims[k] = (long long)le(m + off + 8 * k, 8);
      d->have_space = 1;
    } else if (type == 0x0003) {
      d->cls = m[0] & 0x0f;
      d->size = (int)le(m + 4, 4);
      if (m[1] & 1) return fail("big-endian datatype");
      d->is_signed = (m[1] >> 3) & 1;
      if (!((d->cls == 0 && (d->size == 1 || d->size == 2 || d->size == 4 || d->size == 8)) ||
            (d->cls == 1 && (d->size == 4 || d->size == 8))))
        return fail("datatype class %d size %d", d->cls, d->size);
      d->have_type = 1;
    } else if (type == 0x0008) {
      if (m[0] != 3) return fail("layout message version %d", m[0]);
      if (m[1] == 1) {
        d->addr = le(m + 2, 8);
        d->bytes = le(m + 10, 8);
      } else if (m[1] == 0) {
        d->bytes = le(m + 2, 2);
        d->compact = m + 4;
      } else {
        return fail("chunked or virtual storage");
      }
      d->have_layout = 1;
    } else if (type == 0x0010) {
      if (parse_messages(f, le(m, 8), le(m + 8, 8), d, depth + 1)) return -1;
    } else if (type == 0x000B || type == 0x000C) {
      if (type == 0x000B) return fail("filtered (compressed) dataset");
    }
    pos += 8 + (uint64_t)size;
  }
  return 0;
}

static int parse_object(const file_t *f, uint64_t oh, dset_t *d) {
  const unsigned char *h;
  if (rd(f, oh, 16, &h)) return fail("object header address out of the file");
  if (h[0] != 1) return fail("object header version %d (only version 1)", h[0]);
  memset(d, 0, sizeof *d);
  return parse_messages(f, oh + 16, le(h + 8, 4), d, 0);
}

/* object header address of `name` in the group whose B-tree / heap are at
 * btree / heap; 0 if absent */
static int find_in_btree(const file_t *f, uint64_t node, const char *heap_data, size_t heap_size, const char *name,
                         uint64_t *found, int depth) {
  const unsigned char *h;
  if (depth > 16 || rd(f, node, 24, &h) || memcmp(h, "TREE", 4) || h[4] != 0) return fail("bad group B-tree node");
  const int level = h[5], used = (int)le(h + 6, 2);
  const unsigned char *kc;
  if (rd(f, node + 24, (size_t)used * 16 + 8, &kc)) return fail("B-tree node out of the file");
  for (int c = 0; c < used; c++) {
    const uint64_t child = le(kc + 8 + 16 * c, 8);
    if (level > 0) {
      if (find_in_btree(f, child, heap_data, heap_size, name, found, depth + 1)) return -1;
      if (*found) return 0;
      continue;
    }
    const unsigned char *s;
    if (rd(f, child, 8, &s) || memcmp(s, "SNOD", 4)) return fail("bad symbol-table node");
    const int ns = (int)le(s + 6, 2);
    const unsigned char *e;
    if (rd(f, child + 8, (size_t)ns * 40, &e)) return fail("symbol-table node out of the file");
    for (int k = 0; k < ns; k++) {
      const uint64_t off = le(e + 40 * k, 8);
      if (off >= heap_size) return fail("name offset outside the heap");
      if (!strncmp(heap_data + off, name, heap_size - off) && strlen(name) < heap_size - off) {
        *found = le(e + 40 * k + 8, 8);
        return 0;
      }
    }
  }
  return 0;
}

static int load(const char *path, file_t *f) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return fail("cannot open %s", path);
  fseek(fp, 0, SEEK_END);
  const long sz = ftell(fp);
  fseek(fp, 0, SEEK_SET);
  unsigned char *p = sz > 0 ? malloc((size_t)sz) : NULL;
  if (!p || fread(p, 1, (size_t)sz, fp) != (size_t)sz) {
    free(p);
    fclose(fp);
    return fail("cannot read %s", path);
  }
  fclose(fp);
  f->p = p;
  f->n = (size_t)sz;
  return 0;
}

static int locate(const file_t *f, const char *name, dset_t *d) {
  const unsigned char *sb;
  if (rd(f, 0, 96, &sb) || memcmp(sb, "\211HDF\r\n\032\n", 8)) return fail("not an HDF5 file (no signature at 0)");
  if (sb[8] != 0 && sb[8] != 1) return fail("superblock version %d (only 0 and 1)", sb[8]);
  if (sb[13] != 8 || sb[14] != 8) return fail("offsets/lengths of %d/%d bytes (only 8)", sb[13], sb[14]);
  const size_t root_at = sb[8] == 0 ? 56 : 60;
  const unsigned char *root;
  if (rd(f, root_at, 40, &root)) return fail("truncated superblock");
  uint64_t btree = 0, heap = 0;
  if (le(root + 16, 4) == 1) {
    btree = le(root + 24, 8);
    heap = le(root + 32, 8);
  } else {  /* symbol-table message of the root object header */
    const unsigned char *h;
    const uint64_t oh = le(root + 8, 8);
    if (rd(f, oh, 16, &h) || h[0] != 1) return fail("root group header");
    uint64_t pos = oh + 16, end = pos + le(h + 8, 4);
    while (pos + 8 <= end) {
      const unsigned char *m;
      if (rd(f, pos, 8, &m)) break;
      const int type = (int)le(m, 2), size = (int)le(m + 2, 2);
      if (type == 0x0011 && !rd(f, pos + 8, 16, &m)) {
        btree = le(m, 8);
        heap = le(m + 8, 8);
        break;
      }
      pos += 8 + (uint64_t)size;
    }
    if (!btree) return fail("root group without a symbol table");
  }
  const unsigned char *hp;
  if (rd(f, heap, 32, &hp) || memcmp(hp, "HEAP", 4)) return fail("bad local heap");
  const uint64_t hsize = le(hp + 8, 8), hdata = le(hp + 24, 8);
  const unsigned char *hd;
  if (rd(f, hdata, (size_t)hsize, &hd)) return fail("heap data out of the file");
  uint64_t oh = 0;
  if (find_in_btree(f, btree, (const char *)hd, (size_t)hsize, name, &oh, 0)) return -1;
  if (!oh) return fail("no dataset %s", name);
  if (parse_object(f, oh, d)) return -1;
  if (!d->have_space || !d->have_type || !d->have_layout) return fail("dataset %s lacks space, type or layout", name);
  return 0;
}

long long emc_h5_read(const char *path, const char *name, int type, void *buf, long long cap, int *rank,
                      long long dims[3]) {
  if (!path || !name || (type != EMC_H5_F32 && type != EMC_H5_I32 && type != EMC_H5_F64))
    return fail("emc_h5_read: invalid arguments");
  file_t f;
  if (load(path, &f)) return -1;
  dset_t d;
  if (locate(&f, name, &d)) {
    free((void *)f.p);
    return -1;
  }
  long long n = 1;
  for (int k = 0; k < d.rank; k++) n *= d.dims[k];
  if (rank) *rank = d.rank;
  if (dims)
    for (int k = 0; k < 3; k++) dims[k] = k < d.rank ? d.dims[k] : 1;
  const unsigned char *src = d.compact;
  if ((uint64_t)n * (uint64_t)d.size > d.bytes || (!src && rd(&f, d.addr, (size_t)n * d.size, &src))) {
    free((void *)f.p);
    return fail("dataset %s: data out of the file", name);
  }
  if (buf) {
    if (n > cap) {
      free((void *)f.p);
      return fail("dataset %s: %lld elements, buffer holds %lld", name, n, cap);
    }
    for (long long i = 0; i < n; i++) {
      const unsigned char *q = src + (size_t)i * d.size;
      double v;
      if (d.cls == 1) {
        if (d.size == 4) {
          float x;
          memcpy(&x, q, 4);
          v = x;
        } else {
          memcpy(&v, q, 8);
        }
      } else {
        uint64_t u = le(q, d.size);
        if (d.is_signed && d.size < 8 && (u >> (8 * d.size - 1)) & 1) u |= ~0ULL << (8 * d.size);
        v = d.is_signed ? (double)(int64_t)u : (double)u;
      }
      if (type == EMC_H5_F32) ((float *)buf)[i] = (float)v;
      else if (type == EMC_H5_F64) ((double *)buf)[i] = v;
      else ((int *)buf)[i] = (int)v;
    }
  }
  free((void *)f.p);
  return n;
}

/* ---- libspimage layout ---- */
int emc_write_spimage(const char *path, const float *real, const int *mask, int nx, int ny, int nz) {
  if (!real || nx <= 0 || ny <= 0 || nz <= 0) return fail("emc_write_spimage: invalid arguments");
  const long long n = (long long)nx * ny * nz;
  int *ones = NULL;
  if (!mask) {
    ones = malloc(sizeof(int) * (size_t)n);
    if (!ones) return fail("emc_write_spimage: out of memory");
    for (long long i = 0; i < n; i++) ones[i] = 1;
    mask = ones;
  }
  const int ndim = nz > 1 ? 3 : 2;
  const float center[3] = {(nx - 1) / 2.0f, (ny - 1) / 2.0f, (nz - 1) / 2.0f};
  const int zero = 0, version = 2;
  const float fzero = 0.0f;
  emc_h5_dataset s[11];
  int k = 0;
  if (ndim == 3) {
    s[k++] = (emc_h5_dataset){"real", EMC_H5_F32, 3, {nz, ny, nx}, real};
    s[k++] = (emc_h5_dataset){"mask", EMC_H5_I32, 3, {nz, ny, nx}, mask};
  } else {
    s[k++] = (emc_h5_dataset){"real", EMC_H5_F32, 2, {ny, nx, 0}, real};
    s[k++] = (emc_h5_dataset){"mask", EMC_H5_I32, 2, {ny, nx, 0}, mask};
  }
  s[k++] = (emc_h5_dataset){"num_dimensions", EMC_H5_I32, 1, {1, 0, 0}, &ndim};
  s[k++] = (emc_h5_dataset){"image_center", EMC_H5_F32, 1, {3, 0, 0}, center};
  s[k++] = (emc_h5_dataset){"phased", EMC_H5_I32, 1, {1, 0, 0}, &zero};
  s[k++] = (emc_h5_dataset){"shifted", EMC_H5_I32, 1, {1, 0, 0}, &zero};
  s[k++] = (emc_h5_dataset){"scaled", EMC_H5_I32, 1, {1, 0, 0}, &zero};
  s[k++] = (emc_h5_dataset){"detector_distance", EMC_H5_F32, 1, {1, 0, 0}, &fzero};
  s[k++] = (emc_h5_dataset){"lambda", EMC_H5_F32, 1, {1, 0, 0}, &fzero};
  s[k++] = (emc_h5_dataset){"pixel_size", EMC_H5_F32, 1, {1, 0, 0}, &fzero};
  s[k++] = (emc_h5_dataset){"version", EMC_H5_I32, 1, {1, 0, 0}, &version};
  const int r = emc_h5_write(path, s, k);
  free(ones);
  return r;
}

int emc_spimage_shape(const char *path, int *nx, int *ny, int *nz) {
  int rank = 0;
  long long dims[3];
  if (emc_h5_read(path, "real", EMC_H5_F32, NULL, 0, &rank, dims) < 0) return -1;
  if (rank == 2) {
    *nz = 1;
    *ny = (int)dims[0];
    *nx = (int)dims[1];
  } else if (rank == 3) {
    *nz = (int)dims[0];
    *ny = (int)dims[1];
    *nx = (int)dims[2];
  } else {
    return fail("%s: \"real\" has rank %d", path, rank);
  }
  return 0;
}

int emc_read_spimage(const char *path, float *real, int *mask, int nx, int ny, int nz) {
  if (!real || nx <= 0 || ny <= 0 || nz <= 0) return fail("emc_read_spimage: invalid arguments");
  const long long n = (long long)nx * ny * nz;
  const long long got = emc_h5_read(path, "real", EMC_H5_F32, real, n, NULL, NULL);
  if (got < 0) return -1;
  if (got != n) return fail("%s: \"real\" holds %lld elements, not %lld", path, got, n);
  float *im = malloc(sizeof(float) * (size_t)n);
  if (!im) return fail("emc_read_spimage: out of memory");
  char saved[sizeof g_err];
  memcpy(saved, g_err, sizeof g_err);
  if (emc_h5_read(path, "imag", EMC_H5_F32, im, n, NULL, NULL) == n)
    for (long long i = 0; i < n; i++) real[i] = hypotf(real[i], im[i]);
  else
    for (long long i = 0; i < n; i++) real[i] = fabsf(real[i]);
  memcpy(g_err, saved, sizeof g_err);
  free(im);
  if (mask) {
    const long long gm = emc_h5_read(path, "mask", EMC_H5_I32, mask, n, NULL, NULL);
    if (gm < 0) return -1;
    if (gm != n) return fail("%s: \"mask\" holds %lld elements, not %lld", path, gm, n);
  }
  return 0;
}

/* ---- the reference driver's ingest (emc.c:777-864) ---- */

/* Detector-sized value at (x, y) of a side x side grid read from a larger
 * image with stride `stride` about its centre: pixel
 * ((int)(stride * ((x - side/2) + 0.5) + nx/2 - 0.5), same for y), the
 * integer arithmetic of emc.c:811-815 (x - side/2 and nx/2 in int, the rest
 * in double through the literal 0.5, truncated to int).  -1 if outside the
 * image. */
static long strided_index(int x, int y, int side, int stride, int nx, int ny) {
  const int sx = (int)(stride * ((float)(x - side / 2) + 0.5) + nx / 2 - 0.5);
  const int sy = (int)(stride * ((float)(y - side / 2) + 0.5) + ny / 2 - 0.5);
  if (sx < 0 || sx >= nx || sy < 0 || sy >= ny) return -1;
  return (long)sy * nx + sx;
}

int emc_read_patterns(const char *pattern, int n_images, int side, int stride, float *images) {
  if (!pattern || n_images <= 0 || side <= 0 || stride <= 0 || !images)
    return fail("emc_read_patterns: invalid arguments");
  char path[4096];
  for (int i = 0; i < n_images; i++) {
    snprintf(path, sizeof path, pattern, i);
    int nx, ny, nz;
    if (emc_spimage_shape(path, &nx, &ny, &nz)) return -1;
    if (nz != 1) return fail("%s: not a 2-D image", path);
    float *img = malloc(sizeof(float) * (size_t)nx * ny);
    if (!img) return fail("emc_read_patterns: out of memory");
    if (emc_read_spimage(path, img, NULL, nx, ny, 1)) {
      free(img);
      return -1;
    }
    float *out = images + (size_t)i * side * side;
    for (int x = 0; x < side; x++)
      for (int y = 0; y < side; y++) {
        const long k = strided_index(x, y, side, stride, nx, ny);
        if (k < 0) {
          free(img);
          return fail("%s: %dx%d image too small for side %d at stride %d", path, nx, ny, side, stride);
        }
        out[y * side + x] = img[k];
      }
    free(img);
  }
  return 0;
}

int emc_read_mask(const char *path, int side, int stride, int *mask) {
  if (!mask || side <= 0 || stride <= 0) return fail("emc_read_mask: invalid arguments");
  if (path) {
    int nx, ny, nz;
    if (emc_spimage_shape(path, &nx, &ny, &nz)) return -1;
    float *img = malloc(sizeof(float) * (size_t)nx * ny * nz);
    if (!img) return fail("emc_read_mask: out of memory");
    if (emc_read_spimage(path, img, NULL, nx, ny, nz)) {
      free(img);
      return -1;
    }
    for (int x = 0; x < side; x++)
      for (int y = 0; y < side; y++) {
        const long k = strided_index(x, y, side, stride, nx, ny);
        if (k < 0) {
          free(img);
          return fail("%s: %dx%d mask too small for side %d at stride %d", path, nx, ny, side, stride);
        }
        mask[y * side + x] = img[k] == 0.0f ? 0 : 1;
      }
    free(img);
  } else {
    for (int p = 0; p < side * side; p++) mask[p] = 1;
  }
  /* circular crop (emc.c:856-864): radius side/2 about (side/2 - 0.5) */
  for (int x = 0; x < side; x++)
    for (int y = 0; y < side; y++)
      if (sqrt(pow((float)x - (float)side / 2.0 + 0.5, 2) + pow((float)y - (float)side / 2.0 + 0.5, 2)) >
          side / 2.0)
        mask[y * side + x] = 0;
  return 0;
}
