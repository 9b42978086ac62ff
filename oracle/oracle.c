/*
 * oracle.c — CPU restatement of hirayaku/Gary's hot path. TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP implementation, loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg. Reference citations are file:line relative to the reference root.
 *
 * Pinning (what checks each function independently of this file):
 *   or_binary_search / _pos / _2phase / or_build_cache: exhaustive sorted lists of size 1..200
 *     against numpy.searchsorted membership for every key in [min-1, max+1] and cache depths
 *     8..128 (the reference's own oracle is std::binary_search, test/SearchTest.cu:224-233).
 *   or_intersect_*: numpy.intersect1d on random and exhaustive tiny lists (the reference's
 *     oracle is std::set_intersection, test/SetOpTest.cu:94-96).
 *   or_csr_from_edges: networkx edge sets; orientation acyclicity and degree-order properties.
 *   or_tri_count: networkx.triangles and trace(A^3)/6 golden fixtures; closed forms (K_n,
 *     stars, bipartite graphs).
 *   or_hom4: brute force over all n^4 vertex maps (n <= 8), networkx clique enumeration,
 *     trace(A^4) by numpy matrix powers.
 *   or_hom4_ranked: equal to or_hom4 on every golden graph and on seeded ER / R-MAT graphs up to
 *     scale 16 (a different algorithm: orientation and wedges instead of the trace).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- Search.cuh ---------------------------------------------------------------------------- */

int or_binary_search(const int32_t *list, int64_t size, int32_t key) {
  /* Search.cuh:48-61 (int indices, l/r inclusive) */
  int64_t l = 0, r = size - 1;
  while (r >= l) {
    int64_t mid = l + (r - l) / 2;
    int32_t v = list[mid];
    if (v == key) return 1;
    if (v < key) l = mid + 1;
    else r = mid - 1;
  }
  return 0;
}

int or_binary_search_pos(const int32_t *list, int64_t size, int32_t key, int32_t *pos) {
  /* Search.cuh:67-81: pos tracks the probe; on exit pos = r */
  int32_t l = 0, r = (int32_t)(size - 1), p = 0;
  while (r >= l) {
    p = l + (r - l) / 2;
    int32_t v = list[p];
    if (v == key) { *pos = p; return 1; }
    if (v < key) l = p + 1;
    else r = p - 1;
  }
  *pos = r;
  return 0;
}

void or_build_cache(const int32_t *list, int64_t size, int32_t *cache, int32_t n_cache) {
  /* Search.cuh:93-100 (64-bit product i*size, as the NOTE at :97 asks) */
  for (int32_t i = 0; i < n_cache; ++i) cache[i] = list[(int64_t)i * size / n_cache];
}

int or_binary_search_2phase(const int32_t *list, int64_t size, const int32_t *cache,
                            int32_t n_cache, int32_t key) {
  /* Search.cuh:120-138 */
  int32_t pos = 0;
  if (or_binary_search_pos(cache, n_cache, key, &pos)) return 1;
  if (pos < 0) return 0;
  int64_t bottom = (int64_t)pos * size / n_cache;
  int64_t top = (int64_t)(pos + 1) * size / n_cache;
  return or_binary_search(list + bottom, top - bottom, key);
}

/* ---- Intersection.cuh ---------------------------------------------------------------------- */

uint64_t or_intersect_count(const int32_t *a, int64_t na, const int32_t *b, int64_t nb) {
  int64_t i = 0, j = 0;
  uint64_t c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (b[j] < a[i]) ++j;
    else { ++c; ++i; ++j; }
  }
  return c;
}

uint64_t or_intersect_bsearch_lanes(const int32_t *a, int64_t na, const int32_t *b, int64_t nb,
                                    int32_t n_cache, int32_t lanes, uint64_t *partials) {
  if (partials)
    for (int32_t l = 0; l < lanes; ++l) partials[l] = 0;
  if (na == 0 || nb == 0) return 0; /* Intersection.cuh:190,197 */
  if (na > nb) { /* swap so that |A| <= |B| (Intersection.cuh:58-60, 120-122) */
    const int32_t *t = a; a = b; b = t;
    int64_t tn = na; na = nb; nb = tn;
  }
  int32_t *cache = NULL;
  if (n_cache > 0) {
    cache = (int32_t *)malloc(sizeof(int32_t) * (size_t)n_cache);
    or_build_cache(b, nb, cache, n_cache);
  }
  uint64_t total = 0;
  for (int32_t l = 0; l < lanes; ++l) {
    uint64_t c = 0;
    for (int64_t k = l; k < na; k += lanes) /* sA.cooperate(warp): lane-strided keys */
      c += n_cache > 0 ? or_binary_search_2phase(b, nb, cache, n_cache, a[k])
                       : or_binary_search(b, nb, a[k]);
    if (partials) partials[l] = c;
    total += c;
  }
  free(cache);
  return total;
}

int64_t or_intersect_emit(const int32_t *a, int64_t na, const int32_t *b, int64_t nb,
                          int32_t *out) {
  int64_t i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (b[j] < a[i]) ++j;
    else { out[c++] = a[i]; ++i; ++j; }
  }
  return c;
}

/* ---- Graph.cpp: COO -> CSR ------------------------------------------------------------------- */

static int cmp_u64(const void *x, const void *y) {
  uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
  return a < b ? -1 : (a > b ? 1 : 0);
}
static int cmp_i32(const void *x, const void *y) {
  int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
  return a < b ? -1 : (a > b ? 1 : 0);
}

int or_csr_from_edges(int32_t n, int64_t ntuples, const int32_t *src, const int32_t *dst,
                      int32_t mode, int64_t *indPtr, int32_t *colIdx, int64_t cap,
                      int64_t *nnz_out) {
  if (n < 0 || ntuples < 0 || mode < 0 || mode > 3) return -1;
  const int ranked = mode == 3;
  if (ranked) mode = 1;
  /* symmetrize (Graph.cpp:19-32) folded into a canonical (lo, hi) key; loops dropped */
  uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(ntuples ? ntuples : 1));
  int64_t k = 0;
  for (int64_t e = 0; e < ntuples; ++e) {
    int32_t u = src[e], v = dst[e];
    if (u < 0 || v < 0 || u >= n || v >= n) { free(keys); return -1; }
    if (u == v) continue;
    uint32_t lo = (uint32_t)(u < v ? u : v), hi = (uint32_t)(u < v ? v : u);
    keys[k++] = ((uint64_t)lo << 32) | hi;
  }
  qsort(keys, (size_t)k, sizeof(uint64_t), cmp_u64);
  int64_t m = 0; /* dedup */
  for (int64_t e = 0; e < k; ++e)
    if (m == 0 || keys[m - 1] != keys[e]) keys[m++] = keys[e];
  int64_t need = mode == 1 ? 2 * m : m;
  if (need > cap) { free(keys); return -1; }
  int64_t *deg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t e = 0; e < m; ++e) {
    deg[keys[e] >> 32]++;
    deg[keys[e] & 0xffffffffu]++;
  }
  if (ranked) { /* relabel: new id = rank of (deg, id); rewrite the keys canonically */
    uint64_t *dk = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    for (int32_t v = 0; v < n; ++v) dk[v] = ((uint64_t)deg[v] << 32) | (uint32_t)v;
    qsort(dk, (size_t)n, sizeof(uint64_t), cmp_u64);
    int32_t *nid = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int64_t *ndeg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int32_t r = 0; r < n; ++r) nid[(uint32_t)(dk[r] & 0xffffffffu)] = r;
    for (int32_t v = 0; v < n; ++v) ndeg[nid[v]] = deg[v];
    for (int64_t e = 0; e < m; ++e) {
      int32_t a = nid[keys[e] >> 32], b = nid[keys[e] & 0xffffffffu];
      keys[e] = a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
    }
    memcpy(deg, ndeg, sizeof(int64_t) * ((size_t)n + 1));
    free(dk); free(nid); free(ndeg);
  }
  int64_t *outdeg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t e = 0; e < m; ++e) {
    int32_t lo = (int32_t)(keys[e] >> 32), hi = (int32_t)(keys[e] & 0xffffffffu);
    if (mode == 1) { outdeg[lo]++; outdeg[hi]++; }
    else if (mode == 2) outdeg[hi]++; /* keep src > dst: hi -> lo */
    else {
      int lo_first = deg[lo] < deg[hi] || (deg[lo] == deg[hi] && lo < hi);
      outdeg[lo_first ? lo : hi]++;
    }
  }
  indPtr[0] = 0;
  for (int32_t v = 0; v < n; ++v) indPtr[v + 1] = indPtr[v] + outdeg[v];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
  memcpy(fill, indPtr, sizeof(int64_t) * ((size_t)n + 1));
  for (int64_t e = 0; e < m; ++e) {
    int32_t lo = (int32_t)(keys[e] >> 32), hi = (int32_t)(keys[e] & 0xffffffffu);
    if (mode == 1) { colIdx[fill[lo]++] = hi; colIdx[fill[hi]++] = lo; }
    else if (mode == 2) colIdx[fill[hi]++] = lo;
    else {
      int lo_first = deg[lo] < deg[hi] || (deg[lo] == deg[hi] && lo < hi);
      if (lo_first) colIdx[fill[lo]++] = hi;
      else colIdx[fill[hi]++] = lo;
    }
  }
  for (int32_t v = 0; v < n; ++v) /* per-row sort (Graph.cpp:80-85) */
    qsort(colIdx + indPtr[v], (size_t)(indPtr[v + 1] - indPtr[v]), sizeof(int32_t), cmp_i32);
  *nnz_out = need;
  free(fill); free(outdeg); free(deg); free(keys);
  return 0;
}

/* ---- TriQueryTest.cu: cpuTriCount ------------------------------------------------------------ */

int32_t or_omp_threads(int32_t threads) {
#ifdef _OPENMP
  int32_t t = 1;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
#pragma omp single
    t = omp_get_num_threads();
  }
  return t;
#else
  (void)threads;
  return 1;
#endif
}

uint64_t or_tri_count(int32_t m, const int64_t *indPtr, const int32_t *colIdx, int32_t threads) {
  return or_tri_count_range(0, m, indPtr, colIdx, threads);
}

uint64_t or_tri_count_range(int32_t vbegin, int32_t vend, const int64_t *indPtr,
                            const int32_t *colIdx, int32_t threads) {
  /* TriQueryTest.cu:17-37 with a 64-bit count */
  uint64_t count = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : count)
  for (int32_t v = vbegin; v < vend; ++v) {
    const int32_t *vAdj = colIdx + indPtr[v];
    int64_t dv = indPtr[v + 1] - indPtr[v];
    uint64_t local = 0;
    for (int64_t i = 0; i < dv; ++i) {
      int32_t u = vAdj[i];
      local += or_intersect_count(colIdx + indPtr[u], indPtr[u + 1] - indPtr[u], vAdj, dv);
    }
    count += local;
  }
  return count;
}

uint64_t or_tri_count_batch(int32_t ngraphs, const int32_t *vbase, const int64_t *indPtr,
                            const int32_t *colIdx, int32_t threads, uint64_t *counts) {
  /* TriQueryTest.cu:17-37 per graph; the loop over graphs is the parallel one */
  uint64_t total = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
  for (int32_t g = 0; g < ngraphs; ++g) {
    uint64_t c = 0;
    for (int32_t v = vbase[g]; v < vbase[g + 1]; ++v) {
      const int32_t *vAdj = colIdx + indPtr[v];
      const int64_t dv = indPtr[v + 1] - indPtr[v];
      for (int64_t i = 0; i < dv; ++i) {
        const int32_t u = vAdj[i];
        c += or_intersect_count(colIdx + indPtr[u], indPtr[u + 1] - indPtr[u], vAdj, dv);
      }
    }
    if (counts) counts[g] = c;
    total += c;
  }
  return total;
}

/* ---- 4-vertex homomorphism counts ------------------------------------------------------------ */

typedef unsigned __int128 u128;

int or_hom4(int32_t n, const int64_t *indPtr, const int32_t *colIdx, int32_t threads,
            or_hom4_result *out) {
  u128 tri = 0, k4 = 0, diam = 0, c4w = 0, sum_c2 = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel
  {
    u128 l_tri = 0, l_k4 = 0, l_diam = 0, l_c4w = 0, l_c2 = 0;
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t *S = NULL;
    int64_t s_cap = 0;
#pragma omp for schedule(dynamic, 16)
    for (int32_t u = 0; u < n; ++u) {
      const int32_t *Nu = colIdx + indPtr[u];
      int64_t du = indPtr[u + 1] - indPtr[u];
      l_c2 += (u128)du * (u128)(du - (du > 0)) / 2;
      /* tr(A^4) row u: codeg(u, w) for every w, including codeg(u, u) = deg(u) */
      int64_t nt = 0;
      for (int64_t i = 0; i < du; ++i) {
        int32_t v = Nu[i];
        const int32_t *Nv = colIdx + indPtr[v];
        int64_t dv = indPtr[v + 1] - indPtr[v];
        for (int64_t j = 0; j < dv; ++j) {
          if (cnt[Nv[j]]++ == 0) touched[nt++] = Nv[j];
        }
        if (v > u) {
          /* edge (u, v) with u < v: t_e and the triangles / 4-cliques above v */
          uint64_t te = or_intersect_count(Nu, du, Nv, dv);
          l_diam += (u128)2 * te * te;
          if (s_cap < du) { s_cap = du; S = (int32_t *)realloc(S, sizeof(int32_t) * (size_t)s_cap); }
          int64_t ns = 0, a = 0, b = 0;
          while (a < du && b < dv) {
            if (Nu[a] < Nv[b]) ++a;
            else if (Nv[b] < Nu[a]) ++b;
            else { if (Nu[a] > v) S[ns++] = Nu[a]; ++a; ++b; }
          }
          l_tri += ns;
          for (int64_t x = 0; x < ns; ++x) {
            const int32_t *Nw = colIdx + indPtr[S[x]];
            int64_t dw = indPtr[S[x] + 1] - indPtr[S[x]];
            l_k4 += or_intersect_count(S + x + 1, ns - x - 1, Nw, dw);
          }
        }
      }
      for (int64_t t = 0; t < nt; ++t) {
        uint64_t c = cnt[touched[t]];
        l_c4w += (u128)c * c;
        cnt[touched[t]] = 0;
      }
    }
#pragma omp critical
    {
      tri += l_tri; k4 += l_k4; diam += l_diam; c4w += l_c4w; sum_c2 += l_c2;
    }
    free(cnt); free(touched); free(S);
  }
  const u128 lim = (u128)UINT64_MAX;
  u128 hk4 = k4 * 24;
  u128 m2 = (u128)indPtr[n]; /* 2m arcs */
  u128 c4num = c4w - 4 * sum_c2 - m2;
  if (tri > lim || hk4 > lim || diam > lim || c4w > lim) return -4;
  out->tri = (uint64_t)tri;
  out->k4 = (uint64_t)k4;
  out->hom_k4 = (uint64_t)hk4;
  out->hom_diamond = (uint64_t)diam;
  out->hom_c4 = (uint64_t)c4w;
  out->c4 = (uint64_t)(c4num / 8);
  return 0;
}

/* ---- 4-vertex counts by orientation (full-size checker) ----------------------------------- */

static int64_t first_above(const int32_t *N, int64_t d, int32_t key) {
  int64_t l = 0, r = d; /* first index with N[i] > key */
  while (l < r) {
    int64_t mid = (l + r) >> 1;
    if (N[mid] <= key) l = mid + 1;
    else r = mid;
  }
  return l;
}

int or_hom4_ranked(int32_t n, const int64_t *indPtr, const int32_t *colIdx, int32_t threads,
                   or_hom4_result *out) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  for (int32_t u = 1; u < n; ++u) /* ranked: degrees non-decreasing in id */
    if (indPtr[u + 1] - indPtr[u] < indPtr[u] - indPtr[u - 1]) return -1;
  /* upper part of each row: ups[u] = first index of N(u) above u; upper-arc ids are
   * ubase[u] + (i - ups[u]) */
  int64_t *ups = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
  int64_t *ubase = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
#pragma omp parallel for schedule(static)
  for (int32_t u = 0; u < n; ++u)
    ups[u] = indPtr[u] + first_above(colIdx + indPtr[u], indPtr[u + 1] - indPtr[u], u);
  ubase[0] = 0;
  for (int32_t u = 0; u < n; ++u) ubase[u + 1] = ubase[u] + (indPtr[u + 1] - ups[u]);
  const int64_t nup = ubase[n];
  uint32_t *t = (uint32_t *)calloc((size_t)nup + 1, sizeof(uint32_t));
  u128 tri = 0, k4 = 0, c4 = 0, sum_c2 = 0, t2 = 0;
#pragma omp parallel
  {
    u128 l_tri = 0, l_k4 = 0, l_c4 = 0, l_c2 = 0;
    int32_t *S = NULL;
    int64_t s_cap = 0;
    uint32_t *cnt = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
#pragma omp for schedule(dynamic, 64)
    for (int32_t u = 0; u < n; ++u) {
      const int64_t du = indPtr[u + 1] - indPtr[u];
      l_c2 += (u128)du * (u128)(du - (du > 0)) / 2;
      const int32_t *Up = colIdx + ups[u];
      const int64_t nu = indPtr[u + 1] - ups[u];
      if (s_cap < nu) {
        s_cap = nu;
        S = (int32_t *)realloc(S, sizeof(int32_t) * (size_t)(s_cap + 1));
      }
      for (int64_t i = 0; i < nu; ++i) {
        const int32_t v = Up[i];
        const int32_t *Vp = colIdx + ups[v];
        const int64_t nv = indPtr[v + 1] - ups[v];
        int64_t a = i + 1, b = 0, ns = 0; /* N+(u) ∩ N+(v): only ids above v */
        while (a < nu && b < nv) {
          if (Up[a] < Vp[b]) ++a;
          else if (Vp[b] < Up[a]) ++b;
          else {
            S[ns++] = Up[a];
#pragma omp atomic
            t[ubase[u] + i]++;
#pragma omp atomic
            t[ubase[u] + a]++;
#pragma omp atomic
            t[ubase[v] + b]++;
            ++a;
            ++b;
          }
        }
        l_tri += ns;
        for (int64_t x = 0; x + 1 < ns; ++x) { /* 4-cliques u < v < S[x] < S[y] */
          const int32_t w = S[x];
          const int32_t *Wp = colIdx + ups[w];
          const int64_t nw = indPtr[w + 1] - ups[w];
          l_k4 += or_intersect_count(S + x + 1, ns - x - 1, Wp, nw);
        }
      }
      /* wedges u - v - w with v, w < u */
      int64_t nt = 0;
      const int32_t *Nu = colIdx + indPtr[u];
      const int64_t nl = ups[u] - indPtr[u];
      for (int64_t i = 0; i < nl; ++i) {
        const int32_t v = Nu[i];
        const int32_t *Nv = colIdx + indPtr[v];
        const int64_t cut = first_above(Nv, indPtr[v + 1] - indPtr[v], u - 1);
        for (int64_t j = 0; j < cut; ++j)
          if (cnt[Nv[j]]++ == 0) touched[nt++] = Nv[j];
      }
      for (int64_t k = 0; k < nt; ++k) {
        const u128 c = cnt[touched[k]];
        l_c4 += c * (c - 1) / 2;
        cnt[touched[k]] = 0;
      }
    }
#pragma omp critical
    {
      tri += l_tri; k4 += l_k4; c4 += l_c4; sum_c2 += l_c2;
    }
    free(S); free(cnt); free(touched);
  }
#pragma omp parallel
  {
    u128 l_t2 = 0;
#pragma omp for schedule(static)
    for (int64_t e = 0; e < nup; ++e) l_t2 += (u128)t[e] * t[e];
#pragma omp critical
    t2 += l_t2;
  }
  free(t); free(ups); free(ubase);
  const u128 lim = (u128)UINT64_MAX;
  const u128 hk4 = k4 * 24, diam = 2 * t2;
  const u128 hc4 = 8 * c4 + 4 * sum_c2 + (u128)indPtr[n];
  if (tri > lim || hk4 > lim || diam > lim || hc4 > lim) return -4;
  out->tri = (uint64_t)tri;
  out->k4 = (uint64_t)k4;
  out->hom_k4 = (uint64_t)hk4;
  out->hom_diamond = (uint64_t)diam;
  out->hom_c4 = (uint64_t)hc4;
  out->c4 = (uint64_t)c4;
  return 0;
}

/* ---- k-cliques --------------------------------------------------------------------------- */

static int64_t kc_merge(const int32_t *a, int64_t na, const int32_t *b, int64_t nb, int32_t *out) {
  int64_t i = 0, j = 0, n = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else {
      out[n++] = a[i];
      ++i;
      ++j;
    }
  }
  return n;
}

/* f(S, j) with S = bufs[level][0..ns); bufs[level + 1] receives the next candidate set. */
static unsigned __int128 kc_rec(const int64_t *indPtr, const int32_t *colIdx, int32_t **bufs,
                                int32_t level, int64_t ns, int32_t j) {
  if (j == 1) return (unsigned __int128)ns;
  const int32_t *S = bufs[level];
  unsigned __int128 acc = 0;
  for (int64_t x = 0; x < ns; ++x) {
    const int32_t v = S[x];
    const int64_t nn = kc_merge(S, ns, colIdx + indPtr[v], indPtr[v + 1] - indPtr[v],
                                bufs[level + 1]);
    if (nn >= j - 1) acc += kc_rec(indPtr, colIdx, bufs, level + 1, nn, j - 1);
  }
  return acc;
}

int or_kclique_count(int32_t m, const int64_t *indPtr, const int32_t *colIdx, int32_t k,
                     int32_t threads, uint64_t *count) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  *count = 0;
  if (k < 1) return -1;
  if (k == 1) {
    *count = (uint64_t)m;
    return 0;
  }
  int64_t maxd = 0;
  for (int32_t v = 0; v < m; ++v)
    if (indPtr[v + 1] - indPtr[v] > maxd) maxd = indPtr[v + 1] - indPtr[v];
  unsigned __int128 total = 0;
#pragma omp parallel reduction(+ : total)
  {
    int32_t **bufs = (int32_t **)malloc(sizeof(int32_t *) * (size_t)k);
    for (int32_t l = 0; l < k; ++l) bufs[l] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(maxd + 1));
#pragma omp for schedule(dynamic, 64)
    for (int32_t u = 0; u < m; ++u) {
      const int64_t d = indPtr[u + 1] - indPtr[u];
      if (d < k - 1) continue;
      memcpy(bufs[0], colIdx + indPtr[u], sizeof(int32_t) * (size_t)d);
      total += kc_rec(indPtr, colIdx, bufs, 0, d, k - 1);
    }
    for (int32_t l = 0; l < k; ++l) free(bufs[l]);
    free(bufs);
  }
  if (total > (unsigned __int128)UINT64_MAX) return -4;
  *count = (uint64_t)total;
  return 0;
}

/* ---- triangle listing ------------------------------------------------------------------- */

uint64_t or_triple_hash(int32_t u, int32_t v, int32_t w) {
  uint64_t x = (uint64_t)(uint32_t)u * 0x9E3779B97F4A7C15ull ^
               (uint64_t)(uint32_t)v * 0xC2B2AE3D27D4EB4Full ^
               (uint64_t)(uint32_t)w * 0x165667B19E3779F9ull;
  x ^= x >> 29;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 32;
  return x;
}

int64_t or_tri_list(int32_t m, const int64_t *indPtr, const int32_t *colIdx, int32_t *out,
                    int64_t cap) {
  int64_t n = 0;
  for (int32_t u = 0; u < m; ++u)
    for (int64_t a = indPtr[u]; a < indPtr[u + 1]; ++a) {
      const int32_t v = colIdx[a];
      int64_t i = indPtr[u], j = indPtr[v];
      while (i < indPtr[u + 1] && j < indPtr[v + 1]) {
        if (colIdx[i] < colIdx[j]) ++i;
        else if (colIdx[i] > colIdx[j]) ++j;
        else {
          if (out) {
            if (n == cap) return n;
            out[3 * n] = u;
            out[3 * n + 1] = v;
            out[3 * n + 2] = colIdx[i];
          }
          ++n;
          ++i;
          ++j;
        }
      }
    }
  return n;
}

uint64_t or_tri_list_checksum(int32_t m, const int64_t *indPtr, const int32_t *colIdx,
                              int32_t threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  uint64_t sum = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : sum)
  for (int32_t u = 0; u < m; ++u)
    for (int64_t a = indPtr[u]; a < indPtr[u + 1]; ++a) {
      const int32_t v = colIdx[a];
      int64_t i = indPtr[u], j = indPtr[v];
      while (i < indPtr[u + 1] && j < indPtr[v + 1]) {
        if (colIdx[i] < colIdx[j]) ++i;
        else if (colIdx[i] > colIdx[j]) ++j;
        else {
          sum += or_triple_hash(u, v, colIdx[i]);
          ++i;
          ++j;
        }
      }
    }
  return sum;
}
