/*
 * CPU oracle for the batched signed-distance path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C restatement of the reference algorithm, in double precision, compiled with
 * -ffp-contract=off so every product and sum rounds as the reference's scalar code does.
 * Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it, and
 * only as the checker. The product never links it.
 *
 * Pinning (each restated function is checked by tests/test_oracle.py):
 *   - pseudo-normals       against the compiled reference's DCEL normals (golden fixtures)
 *   - triangle distance     known answers of Tests/TestTriangle.cpp:48-113, closed-form
 *                           closest point (Ericson 5.1.5) on random triangles
 *   - mesh SDF + closest    golden fixtures from the compiled reference (d and face id up
 *                           to exact ties), brute force, sphere invariants
 *   - bins / Morton / order golden fixtures, encode(3,5,7)=431, decode round trips
 *   - CSG union             golden fixtures from BVHUnionIF, sphere/box closed forms
 */
#define _GNU_SOURCE
#include "ebg_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* small vector helpers (Source/EBGeometry_VecImplem.hpp: dot :485-488, cross :372-375,   */
/* length :492-495, operator/(T) multiplies by the reciprocal :303-309)                    */
/* ------------------------------------------------------------------------------------ */

typedef struct { double x[3]; } v3;

static inline v3 mk(double a, double b, double c) { v3 r = {{a, b, c}}; return r; }
static inline v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]); }
static inline v3 add(v3 a, v3 b) { return mk(a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]); }
static inline v3 scl(double s, v3 a) { return mk(s * a.x[0], s * a.x[1], s * a.x[2]); }
static inline double dot(v3 a, v3 b) { return a.x[0] * b.x[0] + a.x[1] * b.x[1] + a.x[2] * b.x[2]; }
static inline double len(v3 a) { return sqrt(a.x[0] * a.x[0] + a.x[1] * a.x[1] + a.x[2] * a.x[2]); }
static inline v3 cross(v3 a, v3 u)
{
  return mk(a.x[1] * u.x[2] - a.x[2] * u.x[1], a.x[2] * u.x[0] - a.x[0] * u.x[2],
            a.x[0] * u.x[1] - a.x[1] * u.x[0]);
}
static inline v3 divs(v3 a, double s)
{
  const double is = 1.0 / s;
  return mk(is * a.x[0], is * a.x[1], is * a.x[2]);
}

/* ------------------------------------------------------------------------------------ */
/* triangle                                                                               */
/* ------------------------------------------------------------------------------------ */

typedef struct
{
  v3       v[3];  /* positions */
  v3       n;     /* face normal recomputed from positions */
  v3       vn[3]; /* vertex pseudo-normals */
  v3       en[3]; /* edge pseudo-normals, v0->v1, v1->v2, v2->v0 */
  uint32_t meta;  /* face index */
} tri_t;

/* Triangle::computeNormal, Source/EBGeometry_TriangleImplem.hpp:94-106 */
static v3 face_normal(v3 x0, v3 x1, v3 x2)
{
  const v3 n = cross(sub(x2, x0), sub(x2, x1));
  return divs(n, len(n));
}

/* TriangleSoAT::signedDistanceLane, Source/EBGeometry_TriangleSoAImplem.hpp:1199-1293.
 * sgn(x) = x > 0 ? 1 : -1 as in the lane path (:1218). */
static double tri_sdf(const tri_t* t, v3 p)
{
#define SGN(x) (((x) > 0.0) ? 1.0 : -1.0)
  const v3 v21 = sub(t->v[1], t->v[0]);
  const v3 v32 = sub(t->v[2], t->v[1]);
  const v3 v13 = sub(t->v[0], t->v[2]);
  const v3 p1  = sub(p, t->v[0]);
  const v3 p2  = sub(p, t->v[1]);
  const v3 p3  = sub(p, t->v[2]);

  const double ss0 = SGN(dot(cross(v21, t->n), p1));
  const double ss1 = SGN(dot(cross(v32, t->n), p2));
  const double ss2 = SGN(dot(cross(v13, t->n), p3));

  if (fabs(ss0 + ss1 + ss2) >= 2.0) return dot(t->n, p1);

  double r2 = dot(p1, p1);
  v3     rv = p1;
  v3     rn = t->vn[0];

  const double p2d = dot(p2, p2);
  if (p2d < r2) { r2 = p2d; rv = p2; rn = t->vn[1]; }
  const double p3d = dot(p3, p3);
  if (p3d < r2) { r2 = p3d; rv = p3; rn = t->vn[2]; }

  const double t1 = dot(p1, v21) / dot(v21, v21);
  if (t1 > 0.0 && t1 < 1.0) {
    const v3     y1 = sub(p1, scl(t1, v21));
    const double yd = dot(y1, y1);
    if (yd < r2) { r2 = yd; rv = y1; rn = t->en[0]; }
  }
  const double t2 = dot(p2, v32) / dot(v32, v32);
  if (t2 > 0.0 && t2 < 1.0) {
    const v3     y2 = sub(p2, scl(t2, v32));
    const double yd = dot(y2, y2);
    if (yd < r2) { r2 = yd; rv = y2; rn = t->en[1]; }
  }
  const double t3 = dot(p3, v13) / dot(v13, v13);
  if (t3 > 0.0 && t3 < 1.0) {
    const v3     y3 = sub(p3, scl(t3, v13));
    const double yd = dot(y3, y3);
    if (yd < r2) { r2 = yd; rv = y3; rn = t->en[2]; }
  }
  return sqrt(r2) * SGN(dot(rn, rv));
#undef SGN
}

double orc_triangle_sdf(const double* v, const double* fn, const double* vn, const double* en,
                        const double* p)
{
  tri_t t;
  for (int i = 0; i < 3; i++) {
    t.v[i]  = ld(v + 3 * i);
    t.vn[i] = ld(vn + 3 * i);
    t.en[i] = ld(en + 3 * i);
  }
  t.n    = ld(fn);
  t.meta = 0;
  return tri_sdf(&t, ld(p));
}

void orc_triangle_normal(const double* v, double* n)
{
  const v3 r = face_normal(ld(v), ld(v + 3), ld(v + 6));
  n[0]       = r.x[0];
  n[1]       = r.x[1];
  n[2]       = r.x[2];
}

/* ------------------------------------------------------------------------------------ */
/* bounding boxes (Source/EBGeometry_BoundingVolumesImplem.hpp)                           */
/* ------------------------------------------------------------------------------------ */

typedef struct { v3 lo, hi; } box_t;

static box_t box_empty(void)
{
  box_t b = {mk(DBL_MAX, DBL_MAX, DBL_MAX), mk(-DBL_MAX, -DBL_MAX, -DBL_MAX)};
  return b;
}
static void box_grow(box_t* b, const box_t* o)
{
  for (int c = 0; c < 3; c++) {
    b->lo.x[c] = fmin(b->lo.x[c], o->lo.x[c]);
    b->hi.x[c] = fmax(b->hi.x[c], o->hi.x[c]);
  }
}
/* getArea :512-521 */
static double box_area(const box_t* b)
{
  const v3 d = sub(b->hi, b->lo);
  return 2.0 * (d.x[0] * d.x[1] + d.x[1] * d.x[2] + d.x[2] * d.x[0]);
}
/* getDistance2 :479-494 */
static double box_dist2(const box_t* b, v3 p)
{
  double s = 0.0;
  double g[3];
  for (int c = 0; c < 3; c++) {
    const double a = b->lo.x[c] - p.x[c];
    const double e = p.x[c] - b->hi.x[c];
    g[c]           = fmax(0.0, a > e ? a : e);
  }
  s = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  return s;
}

/* ------------------------------------------------------------------------------------ */
/* binned SAH K-ary build (Source/EBGeometry_BVH.hpp:521-734) + DFS packing               */
/* (Source/EBGeometry_BVHImplem.hpp:164-205, 405-548)                                     */
/* ------------------------------------------------------------------------------------ */

typedef struct
{
  uint32_t prim;
  box_t    bv;
  v3       c; /* centroid 0.5*(lo+hi), getCentroid :416-424 */
} item_t;

typedef struct
{
  box_t    bv;
  uint32_t leaf;   /* 1 for a leaf */
  uint32_t off;    /* leaf: first primitive (group) */
  uint32_t cnt;    /* leaf: number of primitives (groups) */
  uint32_t child[16];
} node_t;

typedef struct
{
  node_t* nodes;
  int64_t nnodes, cap;
  int     K;
} tree_t;

static int64_t tree_push(tree_t* t)
{
  if (t->nnodes == t->cap) {
    t->cap   = t->cap ? 2 * t->cap : 1024;
    t->nodes = (node_t*)realloc(t->nodes, (size_t)t->cap * sizeof(node_t));
  }
  memset(&t->nodes[t->nnodes], 0, sizeof(node_t));
  return t->nnodes++;
}

/* SAH2WaySplit :521-641 (stable partition; the reference's std::partition order is
 * unspecified, which only changes the tree shape, never a distance). */
static int64_t sah_2way(item_t* it, int64_t b, int64_t e, item_t* tmp)
{
  enum { BINS = 32 };
  const int64_t N = e - b;
  v3            clo = mk(DBL_MAX, DBL_MAX, DBL_MAX), chi = mk(-DBL_MAX, -DBL_MAX, -DBL_MAX);
  for (int64_t i = b; i < e; i++)
    for (int c = 0; c < 3; c++) {
      clo.x[c] = fmin(clo.x[c], it[i].c.x[c]);
      chi.x[c] = fmax(chi.x[c], it[i].c.x[c]);
    }
  double bestCost = DBL_MAX, bestPlane = 0.0;
  int    bestAxis = -1;
  for (int axis = 0; axis < 3; axis++) {
    const double lo = clo.x[axis], hi = chi.x[axis], ext = hi - lo;
    if (ext <= 0.0) continue;
    const double scale = (double)BINS / ext;
    box_t        bins[BINS];
    int          cnt[BINS];
    double       leftArea[BINS - 1];
    int          leftCnt[BINS - 1];
    for (int k = 0; k < BINS; k++) { bins[k] = box_empty(); cnt[k] = 0; }
    for (int64_t i = b; i < e; i++) {
      int k = (int)((it[i].c.x[axis] - lo) * scale);
      if (k > BINS - 1) k = BINS - 1;
      box_grow(&bins[k], &it[i].bv);
      cnt[k]++;
    }
    box_t r = box_empty();
    int   rc = 0;
    for (int k = 0; k < BINS - 1; k++) {
      if (cnt[k] > 0) box_grow(&r, &bins[k]);
      rc += cnt[k];
      leftArea[k] = rc > 0 ? box_area(&r) : 0.0;
      leftCnt[k]  = rc;
    }
    box_t rr  = box_empty();
    int   rrc = 0;
    for (int k = BINS - 1; k >= 1; k--) {
      if (cnt[k] > 0) box_grow(&rr, &bins[k]);
      rrc += cnt[k];
      if (leftCnt[k - 1] > 0 && rrc > 0) {
        const double cost = leftArea[k - 1] * (double)leftCnt[k - 1] + box_area(&rr) * (double)rrc;
        if (cost < bestCost) {
          bestCost  = cost;
          bestAxis  = axis;
          bestPlane = lo + (double)k / scale;
        }
      }
    }
  }
  if (bestAxis < 0) return b + N / 2;
  int64_t nl = 0, nr = 0;
  for (int64_t i = b; i < e; i++)
    if (it[i].c.x[bestAxis] < bestPlane) it[b + nl++] = it[i];
    else tmp[nr++] = it[i];
  memcpy(it + b + nl, tmp, (size_t)nr * sizeof(item_t));
  const int64_t s = b + nl;
  return (s == b || s == e) ? b + N / 2 : s;
}

/* SAHKWaySplit :657-687 */
static void sah_kway(item_t* it, int64_t b, int64_t e, int K, int64_t* ranges, int* nr, item_t* tmp)
{
  if (K <= 1 || b >= e) {
    ranges[2 * *nr]     = b;
    ranges[2 * *nr + 1] = e;
    (*nr)++;
    return;
  }
  const int     K1  = K / 2, K2 = K - K1;
  const int64_t raw = sah_2way(it, b, e, tmp);
  int64_t       mid = raw;
  if (mid > e - K2) mid = e - K2;
  if (mid < b + K1) mid = b + K1;
  sah_kway(it, b, mid, K1, ranges, nr, tmp);
  sah_kway(it, mid, e, K2, ranges, nr, tmp);
}

typedef struct
{
  tree_t   tree;
  item_t*  items;
  item_t*  tmp;
  int64_t  max_leaf;     /* leaf iff size <= max_leaf (mesh) */
  int      leaf_lt_k;    /* leaf iff size < K (CSG stop criterion) */
  /* leaf callback output: items in leaf order */
  uint32_t* leaf_prims;
  int64_t   nleaf_prims;
  int       W;           /* group width; 1 for CSG */
  int64_t   ngroups;
  uint32_t* group_first; /* first leaf_prims index of each group */
  uint32_t* group_count;
} builder_t;

/* topDownSortAndPartition + PackedBVH DFS packing fused: children are built in order
 * c = 0..K-1, so pre-order indices and leaf order match the two-pass reference. */
static uint32_t build_rec(builder_t* B, int64_t b, int64_t e)
{
  const int64_t idx = tree_push(&B->tree);
  box_t         bv  = box_empty();
  for (int64_t i = b; i < e; i++) box_grow(&bv, &B->items[i].bv);
  B->tree.nodes[idx].bv = bv;
  const int64_t n       = e - b;
  const int     K       = B->tree.K;
  const int     stop    = B->leaf_lt_k ? (n < K) : (n <= B->max_leaf);
  if (stop || n < K) {
    node_t* nd = &B->tree.nodes[idx];
    nd->leaf   = 1;
    nd->off    = (uint32_t)B->ngroups;
    const int64_t ng = (n + B->W - 1) / B->W;
    nd->cnt          = (uint32_t)ng;
    for (int64_t g = 0; g < ng; g++) {
      const int64_t gb = b + g * B->W;
      const int64_t ge = (gb + B->W < e) ? gb + B->W : e;
      B->group_first[B->ngroups] = (uint32_t)B->nleaf_prims;
      B->group_count[B->ngroups] = (uint32_t)(ge - gb);
      B->ngroups++;
      for (int64_t i = gb; i < ge; i++) B->leaf_prims[B->nleaf_prims++] = B->items[i].prim;
    }
    return (uint32_t)idx;
  }
  int64_t ranges[32];
  int     nr = 0;
  sah_kway(B->items, b, e, K, ranges, &nr, B->tmp);
  for (int k = 0; k < K; k++) {
    const uint32_t c             = build_rec(B, ranges[2 * k], ranges[2 * k + 1]);
    B->tree.nodes[idx].child[k] = c;
  }
  return (uint32_t)idx;
}

typedef struct
{
  tree_t    tree;
  uint32_t* prims;   /* leaf order */
  uint32_t* gfirst;
  uint32_t* gcount;
  int64_t   ngroups;
} bvh_t;

static void bvh_build(bvh_t* out, const box_t* bvs, int64_t n, int K, int W, int64_t max_leaf, int leaf_lt_k)
{
  builder_t B;
  memset(&B, 0, sizeof(B));
  B.tree.K      = K;
  B.items       = (item_t*)malloc((size_t)n * sizeof(item_t));
  B.tmp         = (item_t*)malloc((size_t)n * sizeof(item_t));
  B.max_leaf    = max_leaf;
  B.leaf_lt_k   = leaf_lt_k;
  B.W           = W;
  B.leaf_prims  = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  B.group_first = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  B.group_count = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
  for (int64_t i = 0; i < n; i++) {
    B.items[i].prim = (uint32_t)i;
    B.items[i].bv   = bvs[i];
    B.items[i].c    = scl(0.5, add(bvs[i].lo, bvs[i].hi));
  }
  build_rec(&B, 0, n);
  free(B.items);
  free(B.tmp);
  out->tree    = B.tree;
  out->prims   = B.leaf_prims;
  out->gfirst  = B.group_first;
  out->gcount  = B.group_count;
  out->ngroups = B.ngroups;
}

static void bvh_free(bvh_t* b)
{
  free(b->tree.nodes);
  free(b->prims);
  free(b->gfirst);
  free(b->gcount);
}

/* ------------------------------------------------------------------------------------ */
/* mesh                                                                                   */
/* ------------------------------------------------------------------------------------ */

struct orc_mesh
{
  int64_t nv, nt;
  tri_t*  tris; /* face order */
  bvh_t   bvh;
  int     K, W;
};

typedef struct { uint64_t key; uint32_t he; } hekey_t;

static int cmp_hekey(const void* a, const void* b)
{
  const uint64_t x = ((const hekey_t*)a)->key, y = ((const hekey_t*)b)->key;
  return (x > y) - (x < y);
}

orc_mesh* orc_mesh_create(const double* xyz, int64_t nv, const uint32_t* tri, int64_t nt, int K, int W,
                          int max_leaf_groups)
{
  if (nv < 3 || nt < 1 || K < 2 || K > 16 || W < 1 || max_leaf_groups < 1) return NULL;
  for (int64_t i = 0; i < 3 * nt; i++)
    if (tri[i] >= (uint64_t)nv) return NULL;

  orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
  m->nv = nv; m->nt = nt; m->K = K; m->W = W;
  m->tris = (tri_t*)calloc((size_t)nt, sizeof(tri_t));

  /* Face normals (DCEL_FaceImplem.hpp:138-165; identical to Triangle::computeNormal). */
  v3* fnorm = (v3*)malloc((size_t)nt * sizeof(v3));
  for (int64_t f = 0; f < nt; f++) {
    const v3 x0 = ld(xyz + 3 * tri[3 * f]), x1 = ld(xyz + 3 * tri[3 * f + 1]), x2 = ld(xyz + 3 * tri[3 * f + 2]);
    fnorm[f] = face_normal(x0, x1, x2);
  }

  /* Edge pseudo-normals (DCEL_EdgeImplem.hpp:130-165): half-edge i of face f runs
   * v_i -> v_{i+1} (SoupImplem.hpp:192-217); normal = normalise(n_f + n_pairface). */
  hekey_t* keys = (hekey_t*)malloc((size_t)(3 * nt) * sizeof(hekey_t));
  for (int64_t f = 0; f < nt; f++)
    for (int i = 0; i < 3; i++) {
      const uint64_t a = tri[3 * f + i], b = tri[3 * f + (i + 1) % 3];
      keys[3 * f + i].key = a * (uint64_t)nv + b;
      keys[3 * f + i].he  = (uint32_t)(3 * f + i);
    }
  qsort(keys, (size_t)(3 * nt), sizeof(hekey_t), cmp_hekey);
  v3* enorm = (v3*)malloc((size_t)(3 * nt) * sizeof(v3));
  for (int64_t f = 0; f < nt; f++)
    for (int i = 0; i < 3; i++) {
      const uint64_t a = tri[3 * f + i], b = tri[3 * f + (i + 1) % 3];
      hekey_t        q = {b * (uint64_t)nv + a, 0};
      hekey_t*       r = (hekey_t*)bsearch(&q, keys, (size_t)(3 * nt), sizeof(hekey_t), cmp_hekey);
      v3             n = fnorm[f];
      if (r) n = add(n, fnorm[r->he / 3]);
      const double l = len(n);
      enorm[3 * f + i] = (l > DBL_EPSILON) ? divs(n, l) : mk(0, 0, 0);
    }
  free(keys);

  /* Vertex pseudo-normals, angle weighted (DCEL_VertexImplem.hpp:190-300): faces in the
   * order they were attached to the vertex (facet order), alpha = acos(clamp(e1.e2)). */
  v3* vnorm = (v3*)calloc((size_t)nv, sizeof(v3));
  for (int64_t f = 0; f < nt; f++)
    for (int i = 0; i < 3; i++) {
      const uint32_t o  = tri[3 * f + i];
      /* in/out neighbours in half-edge order starting at edge v0->v1 */
      uint32_t       nb[2];
      int            k = 0;
      for (int e = 0; e < 3; e++) {
        const uint32_t a = tri[3 * f + e], b = tri[3 * f + (e + 1) % 3];
        if (a == o) nb[k++] = b;
        else if (b == o) nb[k++] = a;
      }
      const v3 x0 = ld(xyz + 3 * o);
      v3       e1 = sub(ld(xyz + 3 * nb[0]), x0);
      v3       e2 = sub(ld(xyz + 3 * nb[1]), x0);
      e1          = divs(e1, len(e1));
      e2          = divs(e2, len(e2));
      double c    = dot(e1, e2);
      c           = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
      const double alpha = acos(c);
      vnorm[o]           = add(vnorm[o], scl(alpha, fnorm[f]));
    }
  for (int64_t v = 0; v < nv; v++) vnorm[v] = divs(vnorm[v], len(vnorm[v]));

  box_t* bvs = (box_t*)malloc((size_t)nt * sizeof(box_t));
  for (int64_t f = 0; f < nt; f++) {
    tri_t* t = &m->tris[f];
    for (int i = 0; i < 3; i++) {
      t->v[i]  = ld(xyz + 3 * tri[3 * f + i]);
      t->vn[i] = vnorm[tri[3 * f + i]];
      t->en[i] = enorm[3 * f + i];
    }
    t->n    = face_normal(t->v[0], t->v[1], t->v[2]);
    t->meta = (uint32_t)f;
    box_t b = box_empty();
    for (int i = 0; i < 3; i++)
      for (int c = 0; c < 3; c++) {
        b.lo.x[c] = fmin(b.lo.x[c], t->v[i].x[c]);
        b.hi.x[c] = fmax(b.hi.x[c], t->v[i].x[c]);
      }
    bvs[f] = b;
  }
  free(fnorm);
  free(enorm);
  free(vnorm);

  /* TriMeshSDF(mesh, SAH, maxLeafGroups): leaf iff size <= maxLeafGroups*W
   * (MeshDistanceFunctionsImplem.hpp:189-191, 453-459). */
  bvh_build(&m->bvh, bvs, nt, K, W, (int64_t)max_leaf_groups * W, 0);
  free(bvs);
  return m;
}

void orc_mesh_destroy(orc_mesh* m)
{
  if (!m) return;
  bvh_free(&m->bvh);
  free(m->tris);
  free(m);
}

int64_t orc_mesh_num_nodes(const orc_mesh* m) { return m->bvh.tree.nnodes; }

void orc_mesh_normals(const orc_mesh* m, double* fn, double* vn, double* en)
{
  for (int64_t f = 0; f < m->nt; f++) {
    const tri_t* t = &m->tris[f];
    for (int c = 0; c < 3; c++) fn[3 * f + c] = t->n.x[c];
    for (int i = 0; i < 3; i++)
      for (int c = 0; c < 3; c++) {
        vn[9 * f + 3 * i + c] = t->vn[i].x[c];
        en[9 * f + 3 * i + c] = t->en[i].x[c];
      }
  }
}

typedef struct { uint32_t idx; double d2; } entry_t;

/* One query: PackedBVH traverse/pruneTraverse (BVHImplem.hpp:1054-1096, 1581-1605) with the
 * TriMeshSDF::getClosestTriangle leaf evaluator (MeshDistanceFunctionsImplem.hpp:511-549). */
static double mesh_query(const orc_mesh* m, v3 p, uint32_t* meta, int64_t* w_inner, int64_t* w_groups,
                         int64_t* w_tris)
{
  const tree_t* T  = &m->bvh.tree;
  const int     K  = T->K;
  double        st = DBL_MAX; /* state, T::max */
  uint32_t      sm = 0;
  entry_t       stack[1024];
  int           top = 0;
  stack[top].idx    = 0;
  stack[top++].d2   = box_dist2(&T->nodes[0].bv, p);
  while (top > 0) {
    const entry_t e = stack[--top];
    if (!(e.d2 <= st * st)) continue;
    const node_t* nd = &T->nodes[e.idx];
    if (nd->leaf) {
      for (uint32_t g = nd->off; g < nd->off + nd->cnt; g++) {
        /* TriangleAoSoA::signedDistance(p, Meta&) :63-87: first minimum over valid lanes */
        const uint32_t f0 = m->bvh.gfirst[g], c = m->bvh.gcount[g];
        double         best = 0.0, babs = 0.0;
        uint32_t       bm   = 0;
        for (uint32_t l = 0; l < c; l++) {
          const tri_t* t  = &m->tris[m->bvh.prims[f0 + l]];
          const double d  = tri_sdf(t, p);
          const double ad = fabs(d);
          if (l == 0 || ad < babs) { best = d; babs = ad; bm = t->meta; }
        }
        if (w_groups) (*w_groups)++;
        if (w_tris) *w_tris += c;
        if (fabs(best) < fabs(st)) { st = best; sm = bm; }
      }
    }
    else {
      if (w_inner) (*w_inner)++;
      entry_t ch[16];
      for (int k = 0; k < K; k++) {
        ch[k].idx = nd->child[k];
        ch[k].d2  = box_dist2(&T->nodes[nd->child[k]].bv, p);
      }
      /* order by descending d2 (stable on child index for equal keys) */
      for (int i = 1; i < K; i++) {
        entry_t x = ch[i];
        int     j = i - 1;
        while (j >= 0 && ch[j].d2 < x.d2) { ch[j + 1] = ch[j]; j--; }
        ch[j + 1] = x;
      }
      for (int k = 0; k < K; k++) stack[top++] = ch[k];
    }
  }
  if (meta) *meta = sm;
  return st;
}

typedef struct
{
  const void*   obj;
  const double* pts;
  double*       out;
  uint32_t*     tri;
  int64_t       b, e;
  int           mode;
} job_t;

static void* mesh_job(void* arg)
{
  job_t*          j = (job_t*)arg;
  const orc_mesh* m = (const orc_mesh*)j->obj;
  for (int64_t i = j->b; i < j->e; i++) {
    const v3 p = ld(j->pts + 3 * i);
    uint32_t meta = 0;
    double   d;
    if (j->mode == 0) d = mesh_query(m, p, &meta, NULL, NULL, NULL);
    else {
      d = 0.0;
      for (int64_t f = 0; f < m->nt; f++) {
        const double x = tri_sdf(&m->tris[f], p);
        if (f == 0 || fabs(x) < fabs(d)) { d = x; meta = (uint32_t)f; }
      }
    }
    j->out[i] = d;
    if (j->tri) j->tri[i] = meta;
  }
  return NULL;
}

static void run_jobs(void* (*fn)(void*), const void* obj, const double* pts, int64_t n, double* out, uint32_t* tri,
                     int nthreads, int mode)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (n < 256) nthreads = 1;
  pthread_t th[256];
  job_t     jobs[256];
  const int64_t chunk = (n + nthreads - 1) / nthreads;
  int           started = 0;
  for (int t = 0; t < nthreads; t++) {
    jobs[t].obj  = obj;
    jobs[t].pts  = pts;
    jobs[t].out  = out;
    jobs[t].tri  = tri;
    jobs[t].b    = t * chunk < n ? t * chunk : n;
    jobs[t].e    = (t + 1) * chunk < n ? (t + 1) * chunk : n;
    jobs[t].mode = mode;
  }
  if (nthreads == 1) { fn(&jobs[0]); return; }
  for (int t = 0; t < nthreads; t++)
    if (pthread_create(&th[t], NULL, fn, &jobs[t]) == 0) started = t + 1;
    else { fn(&jobs[t]); }
  for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
}

void orc_mesh_sdf(const orc_mesh* m, const double* pts, int64_t n, double* d, uint32_t* tri, int nthreads)
{
  run_jobs(mesh_job, m, pts, n, d, tri, nthreads, 0);
}

void orc_mesh_sdf_brute(const orc_mesh* m, const double* pts, int64_t n, double* d, uint32_t* tri, int nthreads)
{
  run_jobs(mesh_job, m, pts, n, d, tri, nthreads, 1);
}

void orc_mesh_work(const orc_mesh* m, const double* pts, int64_t n, int64_t* inner, int64_t* groups, int64_t* tris)
{
  for (int64_t i = 0; i < n; i++) {
    inner[i] = groups[i] = tris[i] = 0;
    (void)mesh_query(m, ld(pts + 3 * i), NULL, &inner[i], &groups[i], &tris[i]);
  }
}

/* ------------------------------------------------------------------------------------ */
/* SFC (Source/EBGeometry_SFCImplem.hpp)                                                   */
/* ------------------------------------------------------------------------------------ */

#define VALID_SPAN 2097151u /* 2^21 - 1, SFC.hpp:61-66 */

/* computeBins<float> :228-276. delta = (max-min) * (1/ValidSpan) (Vec3/T multiplies by the
 * rounded reciprocal); bins = clamp(floor((p-min)/delta)) with a true division. */
void orc_bins_f32(const float* pts, int64_t n, uint32_t* bins)
{
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      const float x = pts[3 * i + c];
      mn[c]         = x < mn[c] ? x : mn[c];
      mx[c]         = x > mx[c] ? x : mx[c];
    }
  const float is = 1.0f / (float)VALID_SPAN;
  float       delta[3];
  for (int c = 0; c < 3; c++) {
    delta[c] = is * (mx[c] - mn[c]);
    if (delta[c] == 0.0f) delta[c] = 1.0f;
  }
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      const float v = (pts[3 * i + c] - mn[c]) / delta[c];
      const float f = floorf(v);
      bins[3 * i + c] = (uint32_t)(f < 0.0f ? 0.0f : (f > (float)VALID_SPAN ? (float)VALID_SPAN : f));
    }
}

void orc_bins_f64(const double* pts, int64_t n, uint32_t* bins)
{
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      const double x = pts[3 * i + c];
      mn[c]          = x < mn[c] ? x : mn[c];
      mx[c]          = x > mx[c] ? x : mx[c];
    }
  const double is = 1.0 / (double)VALID_SPAN;
  double       delta[3];
  for (int c = 0; c < 3; c++) {
    delta[c] = is * (mx[c] - mn[c]);
    if (delta[c] == 0.0) delta[c] = 1.0;
  }
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      const double v = (pts[3 * i + c] - mn[c]) / delta[c];
      const double f = floor(v);
      bins[3 * i + c] = (uint32_t)(f < 0.0 ? 0.0 : (f > (double)VALID_SPAN ? (double)VALID_SPAN : f));
    }
}

/* Morton::encode :31-58: spread each 21-bit coordinate to every third bit. */
static uint64_t part_by_3(uint64_t a)
{
  static const uint64_t M[6] = {0x1fffffull, 0x1f00000000ffffull, 0x1f0000ff0000ffull,
                                0x100f00f00f00f00full, 0x10c30c30c30c30c3ull, 0x1249249249249249ull};
  uint64_t b = a & M[0];
  b          = (b | b << 32) & M[1];
  b          = (b | b << 16) & M[2];
  b          = (b | b << 8) & M[3];
  b          = (b | b << 4) & M[4];
  b          = (b | b << 2) & M[5];
  return b;
}

void orc_morton_encode(const uint32_t* bins, int64_t n, uint64_t* codes)
{
  for (int64_t i = 0; i < n; i++)
    codes[i] = part_by_3(bins[3 * i]) | part_by_3(bins[3 * i + 1]) << 1 | part_by_3(bins[3 * i + 2]) << 2;
}

/* Morton::decode :61-82 restated as a plain bit gather. */
void orc_morton_decode(const uint64_t* codes, int64_t n, uint32_t* bins)
{
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) {
      uint32_t x = 0;
      for (int b = 0; b < 21; b++) x |= (uint32_t)((codes[i] >> (3 * b + c)) & 1u) << b;
      bins[3 * i + c] = x;
    }
}

static const uint64_t* g_sort_codes;
static int cmp_code_idx(const void* a, const void* b)
{
  const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  const uint64_t cx = g_sort_codes[x], cy = g_sort_codes[y];
  if (cx != cy) return cx < cy ? -1 : 1;
  return (x > y) - (x < y);
}

void orc_sort_codes(const uint64_t* codes, int64_t n, uint32_t* perm)
{
  static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  for (int64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  pthread_mutex_lock(&mu);
  g_sort_codes = codes;
  qsort(perm, (size_t)n, sizeof(uint32_t), cmp_code_idx);
  pthread_mutex_unlock(&mu);
}

/* ------------------------------------------------------------------------------------ */
/* CSG union                                                                             */
/* ------------------------------------------------------------------------------------ */

typedef struct
{
  const orc_mesh* mesh;
  v3              t;
  double          c[3], s[3]; /* cos/sin of the x, y, z rotation angles */
  double          scale;
} inst_t;

struct orc_scene
{
  int64_t  ns, nb, nm;
  double*  sc;
  double*  sr;
  double*  blo;
  double*  bhi;
  inst_t*  inst;
};

orc_scene* orc_scene_create(int64_t ns, const double* sc, const double* sr, int64_t nb, const double* blo,
                            const double* bhi, int64_t nm, const orc_mesh* const* meshes, const double* mt,
                            const double* mrot, const double* ms)
{
  orc_scene* s = (orc_scene*)calloc(1, sizeof(orc_scene));
  s->ns = ns; s->nb = nb; s->nm = nm;
  s->sc   = (double*)malloc((size_t)(3 * ns + 1) * sizeof(double));
  s->sr   = (double*)malloc((size_t)(ns + 1) * sizeof(double));
  s->blo  = (double*)malloc((size_t)(3 * nb + 1) * sizeof(double));
  s->bhi  = (double*)malloc((size_t)(3 * nb + 1) * sizeof(double));
  s->inst = (inst_t*)calloc((size_t)(nm + 1), sizeof(inst_t));
  if (ns) { memcpy(s->sc, sc, (size_t)(3 * ns) * sizeof(double)); memcpy(s->sr, sr, (size_t)ns * sizeof(double)); }
  if (nb) { memcpy(s->blo, blo, (size_t)(3 * nb) * sizeof(double)); memcpy(s->bhi, bhi, (size_t)(3 * nb) * sizeof(double)); }
  const double pi = 3.14159265358979323846; /* pi<T>, Constants.hpp:23 */
  for (int64_t i = 0; i < nm; i++) {
    s->inst[i].mesh  = meshes[i];
    s->inst[i].t     = ld(mt + 3 * i);
    s->inst[i].scale = ms[i];
    for (int a = 0; a < 3; a++) {
      /* RotateIF ctor :164-183: theta = angle * pi / 180 */
      const double th  = mrot[3 * i + a] * pi / 180.0;
      s->inst[i].c[a] = cos(th);
      s->inst[i].s[a] = sin(th);
    }
  }
  return s;
}

void orc_scene_destroy(orc_scene* s)
{
  if (!s) return;
  free(s->sc); free(s->sr); free(s->blo); free(s->bhi); free(s->inst);
  free(s);
}

/* RotateIF::value :185-225 */
static v3 rotate_point(v3 p, double c, double s, int axis)
{
  const double x = p.x[0], y = p.x[1], z = p.x[2];
  v3           r = p;
  if (axis == 0) { r.x[1] = y * c + z * s; r.x[2] = -y * s + z * c; }
  else if (axis == 1) { r.x[0] = x * c - z * s; r.x[2] = x * s + z * c; }
  else { r.x[0] = x * c + y * s; r.x[1] = -x * s + y * c; }
  return r;
}

static double scene_value(const orc_scene* s, v3 p)
{
  double v = INFINITY;
  for (int64_t i = 0; i < s->ns; i++) {
    /* SphereSDF::signedDistance :240 */
    const double x = len(sub(p, ld(s->sc + 3 * i))) - s->sr[i];
    v              = x < v ? x : v;
  }
  for (int64_t i = 0; i < s->nb; i++) {
    /* BoxSDF::signedDistance :377-... */
    double dl[3];
    for (int c = 0; c < 3; c++) {
      const double a = s->blo[3 * i + c] - p.x[c], b = p.x[c] - s->bhi[3 * i + c];
      dl[c]          = a > b ? a : b;
    }
    int md = 0;
    for (int c = 0; c < 3; c++)
      if (dl[c] > dl[md]) md = c;
    const double in = dl[md] < 0.0 ? dl[md] : 0.0;
    const v3     o  = mk(fmax(0.0, dl[0]), fmax(0.0, dl[1]), fmax(0.0, dl[2]));
    const double x  = in + len(o);
    v               = x < v ? x : v;
  }
  for (int64_t i = 0; i < s->nm; i++) {
    const inst_t* I = &s->inst[i];
    v3            q = sub(p, I->t);                 /* TranslateIF :154-162 */
    q               = rotate_point(q, I->c[2], I->s[2], 2);
    q               = rotate_point(q, I->c[1], I->s[1], 1);
    q               = rotate_point(q, I->c[0], I->s[0], 0);
    q               = divs(q, I->scale);            /* ScaleIF :262-270 */
    const double x  = mesh_query(I->mesh, q, NULL, NULL, NULL, NULL) * I->scale;
    v               = x < v ? x : v;
  }
  return v;
}

static void* scene_job(void* arg)
{
  job_t*           j = (job_t*)arg;
  const orc_scene* s = (const orc_scene*)j->obj;
  for (int64_t i = j->b; i < j->e; i++) j->out[i] = scene_value(s, ld(j->pts + 3 * i));
  return NULL;
}

void orc_scene_value(const orc_scene* s, const double* pts, int64_t n, double* v, int nthreads)
{
  run_jobs(scene_job, s, pts, n, v, NULL, nthreads, 0);
}
