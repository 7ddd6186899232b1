/*
 * Rotation sampling table (emc.c:40-363, generate_rotation_list).
 *
 * The sampling is the 600-cell subdivision of the unit 3-sphere: the 120
 * vertices of the 600-cell plus, for sampling n, the interior lattice points
 * of every edge, triangular face and tetrahedral cell subdivided n times.
 * The reference fixes the order of the entries and the float arithmetic that
 * produces them; both are reproduced here because the table is an input of
 * the GPU path and slices are indexed by table position:
 *   - vertices in the order of emc.c:57-103;
 *   - simplices discovered as in emc.c:105-210 (edges |vi+vj|^2 > 3,
 *     faces > 7.5, cells > 13.5, with its "done" bookkeeping);
 *   - interior points with barycentric coefficients (n - k_m)/n on the
 *     simplex vertices, evaluated in float, left to right (emc.c:242-334);
 *   - weights f_d / N / |p|^3 of the un-normalised point, normalised by a
 *     float running sum, then the points normalised (emc.c:212-360).
 *
 * The half table (one entry per rotation) pairs q with -q combinatorially:
 * the antipode of a lattice point on simplex s is the point with the same
 * barycentric coordinates on the antipodal simplex -s, so no floating-point
 * tolerance is involved.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "emc_host.h"

#define NV 120
#define NE 720
#define NF 1200
#define NC 600

int emc_n_to_samples(int n) { return 20 * (n + 5 * n * n * n); }

/* Golden ratio with the reference's truncated sqrt(5) (emc.c:41). */
static float golden(void) { return (float)((1.0 + 2.23606798) / 2.0); }

static void build_vertices(float v[NV][4]) {
  memset(v, 0, sizeof(float) * NV * 4);
  /* (+-1/2, +-1/2, +-1/2, +-1/2), bit i of the index selects the sign */
  for (int b = 0; b < 16; b++)
    for (int k = 0; k < 4; k++) v[b][k] = (float)(-0.5 + (float)((b >> (3 - k)) & 1));
  /* (+-1, 0, 0, 0) and permutations */
  for (int i = 0; i < 8; i++) v[16 + i][i / 2] = (float)(-1.0 + 2.0 * (float)(i % 2));
  /* even permutations of (+-1/2, +-tau/2, +-1/(2 tau), 0): position of each
   * of the four values, one row per permutation */
  static const unsigned char slot[12][4] = {{0, 1, 2, 3}, {0, 3, 1, 2}, {0, 2, 3, 1}, {1, 2, 0, 3},
                                            {1, 3, 2, 0}, {1, 0, 3, 2}, {2, 0, 1, 3}, {2, 3, 0, 1},
                                            {2, 1, 3, 0}, {3, 1, 0, 2}, {3, 2, 1, 0}, {3, 0, 2, 1}};
  const double t = (double)golden();
  for (int pmt = 0; pmt < 12; pmt++)
    for (int s = 0; s < 8; s++) {
      float *q = v[24 + 8 * pmt + s];
      double h1 = -0.5 + 1.0 * (float)((s >> 2) & 1);
      double h2 = -0.5 + 1.0 * (float)((s >> 1) & 1);
      double h3 = -0.5 + 1.0 * (float)(s & 1);
      q[slot[pmt][0]] = (float)h1;
      q[slot[pmt][1]] = (float)(t * h2);
      q[slot[pmt][2]] = (float)(1.0 / t * h3);
      q[slot[pmt][3]] = 0.0f;
    }
}

/* |sum of the listed vertices|^2: components summed in float in list order,
 * squares summed in double, result rounded to float (emc.c:115-119). */
static float norm2_of_sum(float v[NV][4], const int *ids, int m) {
  double acc = 0.0;
  for (int k = 0; k < 4; k++) {
    float s = v[ids[0]][k];
    for (int j = 1; j < m; j++) s += v[ids[j]][k];
    acc += (double)s * (double)s;
  }
  return (float)acc;
}

typedef struct {
  float v[NV][4];
  int edge[NE][2];
  int face[NF][3];
  int cell[NC][4];
} polytope;

static int build_polytope(polytope *P) {
  build_vertices(P->v);
  int ne = 0;
  for (int i = 0; i < NV; i++)
    for (int j = 0; j < i; j++) {
      int ids[2] = {i, j};
      if (norm2_of_sum(P->v, ids, 2) > 3.0f) {
        if (ne == NE) return -1;
        P->edge[ne][0] = i;
        P->edge[ne][1] = j;
        ne++;
      }
    }
  if (ne != NE) return -1;
  /* a face is reported once, with the first edge (in edge order) that
   * contains two of its vertices, the third vertex not yet seen in an edge */
  unsigned char seen[NV] = {0};
  int nf = 0;
  for (int e = 0; e < NE; e++) {
    seen[P->edge[e][0]] = seen[P->edge[e][1]] = 1;
    for (int j = 0; j < NV; j++) {
      if (seen[j]) continue;
      int ids[3] = {j, P->edge[e][0], P->edge[e][1]};
      if (norm2_of_sum(P->v, ids, 3) > 7.5f) {
        if (nf == NF) return -1;
        P->face[nf][0] = P->edge[e][0];
        P->face[nf][1] = P->edge[e][1];
        P->face[nf][2] = j;
        nf++;
      }
    }
  }
  if (nf != NF) return -1;
  /* a cell is reported once, at its lowest-numbered vertex */
  unsigned char used[NV] = {0};
  int nc = 0;
  for (int j = 0; j < NV; j++) {
    used[j] = 1;
    for (int f = 0; f < NF; f++) {
      if (used[P->face[f][0]] || used[P->face[f][1]] || used[P->face[f][2]]) continue;
      int ids[4] = {P->face[f][0], P->face[f][1], P->face[f][2], j};
      if (norm2_of_sum(P->v, ids, 4) > 13.5f) {
        if (nc == NC) return -1;
        for (int k = 0; k < 3; k++) P->cell[nc][k] = P->face[f][k];
        P->cell[nc][3] = j;
        nc++;
      }
    }
  }
  return nc == NC ? 0 : -1;
}

/* Interior lattice of a d-simplex subdivided n times: tuples (k_0..k_d) with
 * d <= k_m <= n-1 and sum k_m = d*n; the point is sum_m (n-k_m)/n * V_m.
 * Order: for d = 1 by increasing k_1, for d = 2, 3 lexicographic in
 * (k_0, ..., k_{d-1}), the order of emc.c:247-248, 270-272 and 306-310. */
typedef struct {
  int d, n, count;
  int *k;     /* count x (d+1) */
  int *rank;  /* n^(d+1) entries: tuple code -> position, -1 if absent */
} lattice;

static int tuple_code(const int *k, int d, int n) {
  int c = 0;
  for (int m = d; m >= 0; m--) c = c * n + k[m];
  return c;
}

static int lattice_build(lattice *L, int d, int n) {
  L->d = d;
  L->n = n;
  L->count = 0;
  L->k = NULL;
  L->rank = NULL;
  if (n - 1 < d) return 0;
  size_t codes = 1;
  for (int m = 0; m <= d; m++) codes *= (size_t)n;
  L->rank = (int *)malloc(sizeof(int) * codes);
  if (!L->rank) return -1;
  for (size_t c = 0; c < codes; c++) L->rank[c] = -1;
  int cap = 1;
  for (int m = 0; m < d; m++) cap *= n;
  L->k = (int *)malloc(sizeof(int) * (size_t)cap * (d + 1));
  if (!L->k) return -1;
  int t[4];
  if (d == 1) {
    for (t[1] = 1; t[1] <= n - 1; t[1]++) {
      t[0] = n - t[1];
      memcpy(&L->k[L->count * 2], t, sizeof(int) * 2);
      L->rank[tuple_code(t, 1, n)] = L->count++;
    }
    return 0;
  }
  /* odometer over (k_0..k_{d-1}) in [d, n-1], last index fastest */
  for (int m = 0; m < d; m++) t[m] = d;
  for (;;) {
    int s = 0;
    for (int m = 0; m < d; m++) s += t[m];
    t[d] = d * n - s;
    if (t[d] >= d && t[d] <= n - 1) {
      memcpy(&L->k[L->count * (d + 1)], t, sizeof(int) * (d + 1));
      L->rank[tuple_code(t, d, n)] = L->count++;
    }
    int m = d - 1;
    while (m >= 0 && t[m] == n - 1) t[m--] = d;
    if (m < 0) break;
    t[m]++;
  }
  return 0;
}

static void lattice_free(lattice *L) {
  free(L->k);
  free(L->rank);
}

/* Point of the lattice tuple on the simplex with vertex ids `ids`. */
static void lattice_point(float v[NV][4], const int *ids, const int *k, int d, int n, float out[4]) {
  float c[4];
  for (int m = 0; m <= d; m++) c[m] = (float)(n - k[m]) / (float)n;
  for (int j = 0; j < 4; j++) {
    float acc = c[0] * v[ids[0]][j];
    for (int m = 1; m <= d; m++) acc += c[m] * v[ids[m]][j];
    out[j] = acc;
  }
}

/* |p|^3 of the un-normalised point: squares in double (exact for floats),
 * pow(., 3/2), rounded to float (emc.c:235-238). */
static float cube_norm(const float p[4]) {
  double s = 0.0;
  for (int j = 0; j < 4; j++) s += (double)p[j] * (double)p[j];
  return (float)pow(s, (double)(3.0f / 2.0f));
}

/* Antipodal simplex and the vertex permutation that maps it. */
static int find_simplex(const int *list, int count, int m, const int *want, int *perm) {
  for (int s = 0; s < count; s++) {
    const int *c = &list[s * m];
    int ok = 1;
    for (int a = 0; a < m && ok; a++) {
      int hit = -1;
      for (int b = 0; b < m; b++)
        if (c[b] == want[a]) hit = b;
      if (hit < 0) ok = 0; else perm[a] = hit;
    }
    if (ok) return s;
  }
  return -1;
}

static int build_table(int n, int half, float *qout, float *wout) {
  if (n < 1) return -1;
  polytope *P = (polytope *)malloc(sizeof(polytope));
  if (!P || build_polytope(P) != 0) { free(P); return -1; }
  const int ns = emc_n_to_samples(n);
  lattice lat[4];
  for (int d = 1; d <= 3; d++) lattice_build(&lat[d], d, n);
  const int per[4] = {1, lat[1].count, lat[2].count, lat[3].count};
  const int nsimp[4] = {NV, NE, NF, NC};
  int base[4];
  base[0] = 0;
  for (int d = 1; d <= 3; d++) base[d] = base[d - 1] + nsimp[d - 1] * per[d - 1];
  if (base[3] + NC * per[3] != ns) {
    for (int d = 1; d <= 3; d++) lattice_free(&lat[d]);
    free(P);
    return -1;
  }
  const int *simp[4] = {NULL, &P->edge[0][0], &P->face[0][0], &P->cell[0][0]};

  float(*pts)[4] = malloc(sizeof(float) * 4 * (size_t)ns);
  float *w = (float *)malloc(sizeof(float) * ns);
  const float alpha = (float)acos(1.0 / 3.0);
  const float fac[4] = {(float)(20.0 * (3.0 * alpha - M_PI) / 4.0 / M_PI), (float)(5.0 * alpha / 2.0 / M_PI),
                        1.0f, 1.0f};
  for (int i = 0; i < NV; i++) {
    memcpy(pts[i], P->v[i], sizeof(float) * 4);
    w[i] = fac[0] / (float)ns / cube_norm(pts[i]);
  }
  for (int d = 1; d <= 3; d++)
    for (int s = 0; s < nsimp[d]; s++)
      for (int j = 0; j < per[d]; j++) {
        int idx = base[d] + s * per[d] + j;
        lattice_point(P->v, &simp[d][s * (d + 1)], &lat[d].k[j * (d + 1)], d, n, pts[idx]);
        w[idx] = fac[d] / (float)ns / cube_norm(pts[idx]);
      }
  float wsum = 0.0f;
  for (int i = 0; i < ns; i++) wsum += w[i];
  for (int i = 0; i < ns; i++) w[i] /= wsum;
  for (int i = 0; i < ns; i++) {
    double s = 0.0;
    for (int j = 0; j < 4; j++) s += (double)pts[i][j] * (double)pts[i][j];
    float len = (float)sqrt(s);
    for (int j = 0; j < 4; j++) pts[i][j] /= len;
  }

  int result = ns;
  if (!half) {
    memcpy(qout, pts, sizeof(float) * 4 * (size_t)ns);
    memcpy(wout, w, sizeof(float) * ns);
  } else {
    /* vertex antipodes: exact negation */
    int anti[NV];
    for (int i = 0; i < NV; i++) {
      anti[i] = -1;
      for (int j = 0; j < NV; j++)
        if (P->v[j][0] == -P->v[i][0] && P->v[j][1] == -P->v[i][1] && P->v[j][2] == -P->v[i][2] &&
            P->v[j][3] == -P->v[i][3])
          anti[i] = j;
      if (anti[i] < 0) result = -1;
    }
    int *partner = (int *)malloc(sizeof(int) * ns);
    for (int i = 0; i < NV && result > 0; i++) partner[i] = anti[i];
    for (int d = 1; d <= 3 && result > 0; d++)
      for (int s = 0; s < nsimp[d] && result > 0; s++) {
        const int *ids = &simp[d][s * (d + 1)];
        int want[4], perm[4];
        for (int m = 0; m <= d; m++) want[m] = anti[ids[m]];
        int t = find_simplex(simp[d], nsimp[d], d + 1, want, perm);
        if (t < 0) { result = -1; break; }
        for (int j = 0; j < per[d]; j++) {
          int kk[4];
          for (int m = 0; m <= d; m++) kk[perm[m]] = lat[d].k[j * (d + 1) + m];
          int r = lat[d].rank[tuple_code(kk, d, n)];
          if (r < 0) { result = -1; break; }
          partner[base[d] + s * per[d] + j] = base[d] + t * per[d] + r;
        }
      }
    if (result > 0) {
      int m = 0;
      for (int i = 0; i < ns; i++) {
        int p = partner[i];
        if (p == i || partner[p] != i) { result = -1; break; }
        for (int j = 0; j < 4; j++)
          if (fabsf(pts[i][j] + pts[p][j]) > 1e-6f) result = -1;
        if (result < 0) break;
        if (p < i) continue; /* the earlier entry of the pair represents it */
        memcpy(&qout[4 * m], pts[i], sizeof(float) * 4);
        wout[m] = (float)((double)w[i] + (double)w[p]);
        m++;
      }
      if (result > 0) result = m;
    }
    free(partner);
  }
  free(pts);
  free(w);
  for (int d = 1; d <= 3; d++) lattice_free(&lat[d]);
  free(P);
  return result;
}

int emc_rotation_table(int n, int half, float *q, float *w) { return build_table(n, half, q, w); }

int generate_rotation_list(const int n, Quaternion ***return_list, real **return_weights) {
  int ns = emc_n_to_samples(n);
  float *q = (float *)malloc(sizeof(float) * 4 * (size_t)ns);
  real *w = (real *)malloc(sizeof(real) * ns);
  if (!q || !w || build_table(n, 0, q, w) != ns) {
    free(q);
    free(w);
    return -1;
  }
  Quaternion **list = (Quaternion **)malloc(sizeof(Quaternion *) * ns);
  for (int i = 0; i < ns; i++) {
    list[i] = (Quaternion *)malloc(sizeof(Quaternion));
    memcpy(list[i]->q, &q[4 * i], sizeof(float) * 4);
  }
  free(q);
  *return_list = list;
  *return_weights = w;
  return ns;
}

int emc_assign_cells(const float *q_fine, int n_fine, const float *q_coarse, int n_coarse, int *cell) {
  if (!q_fine || !q_coarse || !cell || n_fine < 0 || n_coarse <= 0) return -1;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n_fine; r++) {
    const float *a = &q_fine[4 * (size_t)r];
    double best = -1.0;
    int bi = 0;
    for (int c = 0; c < n_coarse; c++) {
      const float *b = &q_coarse[4 * (size_t)c];
      const double d = fabs((double)a[0] * b[0] + (double)a[1] * b[1] + (double)a[2] * b[2] + (double)a[3] * b[3]);
      if (d > best) {
        best = d;
        bi = c;
      }
    }
    cell[r] = bi;
  }
  return 0;
}
