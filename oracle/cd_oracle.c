/*
 * ORACLE -- test infrastructure only. CPU restatement of the reference's
 * convection-diffusion assemble-and-solve path, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker and
 * baseline; never part of the shipped compute path.
 *
 * It follows the reference's algorithms and operation order, in code of its
 * own:
 *   mesh + regular refinement, depth-first cell order: src/Geometry/Tests.C,
 *     src/Refinement/Ref{Tri,Quad,Tetra,Hexa}RegDesc.C, It_EQ.C:16-51
 *   DOF numbering by union-find over shared joints + renumbering:
 *     src/FE/FESpace3D.C:186-804, FEMapper.C:456-600, FESpace.C:1863-2099
 *   sparsity: src/Matrix/Structure.C:226-618
 *   assembly: Assemble3D.C:183-549, Assemble2D.C:2798-3255,
 *     LocalAssembling.C:734-970, ConvDiff.C:269-340,1776-1999,
 *     HexaAffin.C/TetraAffin.C/TriaAffin.C/QuadAffin.C
 *   BlockFEMatrix::apply_scaled_add: BlockFEMatrix.C:1008-1095, FEMatrix.C:209-223
 *   CG: Iteration_cg.C:27-147; FGMRES: Iteration_gmres.C:430-620,
 *     Utilities_gmres.C:6-108; stopping test: IterativeMethod.h:83-162
 *   multigrid V-cycle: Multigrid.C:157-410, MultigridLevel.C:118-151,
 *     JacobiSmoother.C:22-30, GridTransfer.C:603-1115 (cell-wise
 *     prolongation/restriction with multiplicity averaging)
 *   errors: FEFunction3D.C:60-248, ConvDiff.C:2124-2183
 * Serial sums everywhere (Blas1.C:21-98). Built with -ffp-contract=off.
 * The coarse/direct solve uses dense LU with partial pivoting in place of
 * UMFPACK (not present here); both solve the same system exactly up to
 * rounding.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXNL 27
#define MAXQ 64

/* ----------------------------------------------------------------------- */
/* shapes                                                                   */
/* ----------------------------------------------------------------------- */
enum { TRI = 0, QUAD = 1, TET = 2, HEX = 3 };
static const int NV[4] = {3, 4, 4, 8};
static const int NJ[4] = {3, 4, 4, 6};
static const int DIM[4] = {2, 2, 3, 3};
/* joint vertex lists (Triangle.C/Quadrangle.C edges, Tetrahedron.C:18-19,
   Hexahedron.C:20-22) */
static const int JV_TRI[3][2] = {{0, 1}, {1, 2}, {2, 0}};
static const int JV_QUAD[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
static const int JV_TET[4][3] = {{0, 1, 2}, {0, 3, 1}, {2, 1, 3}, {0, 2, 3}};
static const int JV_HEX[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                                 {2, 6, 7, 3}, {0, 3, 7, 4}, {4, 7, 6, 5}};
static int jv_len(int s) { return s == TRI || s == QUAD ? 2 : (s == TET ? 3 : 4); }
static int jv(int s, int j, int k) {
  switch(s) {
    case TRI: return k < 2 ? JV_TRI[j][k] : -1; /* edges have 2 vertices */
    case QUAD: return k < 2 ? JV_QUAD[j][k] : -1;
    case TET: return JV_TET[j][k];
    default: return JV_HEX[j][k];
  }
}

/* Planted mistakes for the tests that show a check can fail (0: none).
   1: the SUPG -eps*Laplacian term dropped; 2: the mixed second reference
   derivatives dropped from the physical Hessian; 3: the Jacobian of the
   AFC Newton scheme dropped; 4: the Anderson mixing skipped; 5: the
   downwind branch of the Kuzmin Jacobian compares E_kj unrounded; 6: the
   hanging coupling DOFs that hang themselves receive nothing; 7: hanging
   rows added in list order instead of the sorted order; 8: the 1-irregular
   closure of the refinement by indicator skipped; 9: the hanging-DOF
   coupling of the assembled system skipped; 11: the conforming closure
   refines only triangles with three refined edges.
   Set only by tests, through orc_set_fault. */
static int orc_fault = 0;
void orc_set_fault(int code) { orc_fault = code; }

/* ----------------------------------------------------------------------- */
/* mesh                                                                     */
/* ----------------------------------------------------------------------- */
typedef struct {
  int shape, ncells, nverts, cap;
  double* X;   /* 3 per vertex */
  int* cv;     /* NV per cell */
  int* parent; /* cell in the coarser mesh */
} omesh;

/* refined-vertex dictionary keyed by the sorted ids of the parent vertices
   spanning the new vertex: cells share a new vertex through their common
   vertex, edge or face (the refinement's topology), not through equal
   coordinates */
typedef struct {
  int cap;
  int* slot;
  int (*key)[4];
} vdict;

static uint64_t hash4(const int* k) {
  uint64_t h = 1469598103934665603ull;
  for(int i = 0; i < 4; ++i) {
    h ^= (uint64_t)(uint32_t)k[i];
    h *= 1099511628211ull;
    h ^= h >> 29;
  }
  return h;
}

static int vertex_id(omesh* m, vdict* d, const int* key, const double* x) {
  uint64_t h = hash4(key) & (uint64_t)(d->cap - 1);
  for(;;) {
    int v = d->slot[h];
    if(v < 0) break;
    if(memcmp(d->key[h], key, sizeof(int) * 4) == 0) return v;
    h = (h + 1) & (uint64_t)(d->cap - 1);
  }
  if(m->nverts == m->cap) {
    m->cap = m->cap * 2 + 16;
    m->X = realloc(m->X, sizeof(double) * 3 * m->cap);
  }
  int v = m->nverts++;
  memcpy(&m->X[3 * v], x, 3 * sizeof(double));
  d->slot[h] = v;
  memcpy(d->key[h], key, sizeof(int) * 4);
  return v;
}

/* id stored for `key`, or -1 */
static int vertex_find(const vdict* d, const int* key) {
  uint64_t h = hash4(key) & (uint64_t)(d->cap - 1);
  for(;;) {
    int v = d->slot[h];
    if(v < 0) return -1;
    if(memcmp(d->key[h], key, sizeof(int) * 4) == 0) return v;
    h = (h + 1) & (uint64_t)(d->cap - 1);
  }
}

static void mesh_free(omesh* m) {
  free(m->X);
  free(m->cv);
  free(m->parent);
}

static void default_mesh(int geometry, omesh* m) {
  static const double sq[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
  static const double cube[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  memset(m, 0, sizeof(*m));
  const int* cells = NULL;
  static const int c_sq[4] = {0, 1, 2, 3};
  static const int c_tt[6] = {0, 1, 2, 0, 2, 3};
  static const int c_tt2[6] = {0, 1, 3, 1, 2, 3};
  static const int c_hex[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  static const int c_tet[24] = {0, 1, 2, 6, 0, 5, 1, 6, 0, 2, 3, 6, 0, 3, 7, 6, 0, 4, 5, 6, 0, 7, 4, 6};
  int nc = 0, nvx = 0;
  switch(geometry) {
    case 0: m->shape = QUAD; cells = c_sq; nc = 1; nvx = 4; break;
    case 1: m->shape = TRI; cells = c_tt; nc = 2; nvx = 4; break;
    case 2: m->shape = TRI; cells = c_tt2; nc = 2; nvx = 4; break;
    case 3: m->shape = HEX; cells = c_hex; nc = 1; nvx = 8; break;
    default: m->shape = TET; cells = c_tet; nc = 6; nvx = 8; break;
  }
  m->nverts = nvx;
  m->cap = nvx;
  m->X = malloc(sizeof(double) * 3 * nvx);
  for(int v = 0; v < nvx; ++v)
    for(int k = 0; k < 3; ++k) m->X[3 * v + k] = nvx == 4 ? sq[v][k] : cube[v][k];
  m->ncells = nc;
  m->cv = malloc(sizeof(int) * nc * NV[m->shape]);
  memcpy(m->cv, cells, sizeof(int) * nc * NV[m->shape]);
  m->parent = malloc(sizeof(int) * nc);
  for(int c = 0; c < nc; ++c) m->parent[c] = -1;
}

/* structured unit cube: (n+1)^3 grid vertices (x fastest), cubes in the same
   order, each a hexahedron or split into the 6 tetrahedra of the default
   tetrahedral cube (cells around the 0-6 diagonal) */
static void structured_mesh(int n, int tetra, omesh* m) {
  static const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  static const int c_tet[6][4] = {{0, 1, 2, 6}, {0, 5, 1, 6}, {0, 2, 3, 6},
                                  {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 7, 4, 6}};
  memset(m, 0, sizeof(*m));
  m->shape = tetra ? TET : HEX;
  const int n1 = n + 1, nv = NV[m->shape];
  m->nverts = n1 * n1 * n1;
  m->cap = m->nverts;
  m->X = malloc(sizeof(double) * 3 * m->nverts);
  for(int k = 0; k < n1; ++k)
    for(int j = 0; j < n1; ++j)
      for(int i = 0; i < n1; ++i) {
        const int v = i + n1 * (j + n1 * k);
        m->X[3 * v] = (double)i / n;
        m->X[3 * v + 1] = (double)j / n;
        m->X[3 * v + 2] = (double)k / n;
      }
  m->ncells = n * n * n * (tetra ? 6 : 1);
  m->cv = malloc(sizeof(int) * m->ncells * nv);
  int c = 0;
  for(int k = 0; k < n; ++k)
    for(int j = 0; j < n; ++j)
      for(int i = 0; i < n; ++i) {
        int g[8];
        for(int a = 0; a < 8; ++a) g[a] = (i + co[a][0]) + n1 * ((j + co[a][1]) + n1 * (k + co[a][2]));
        if(tetra)
          for(int t = 0; t < 6; ++t, ++c)
            for(int a = 0; a < 4; ++a) m->cv[c * 4 + a] = g[c_tet[t][a]];
        else {
          for(int a = 0; a < 8; ++a) m->cv[c * 8 + a] = g[a];
          ++c;
        }
      }
  m->parent = malloc(sizeof(int) * m->ncells);
  for(int q = 0; q < m->ncells; ++q) m->parent[q] = -1;
}

/* child vertex lists in terms of the parent's refined vertices; refined
   vertex r is the barycentre of the parent vertices listed in span */
typedef struct {
  int nref, nchild;
  int span[27][9]; /* span[r][0] = count */
  int child[8][8];
} reftab;

static void set_span(reftab* t, int r, int n, const int* v) {
  t->span[r][0] = n;
  for(int i = 0; i < n; ++i) t->span[r][1 + i] = v[i];
}

static reftab make_reftab(int shape) {
  reftab t;
  memset(&t, 0, sizeof(t));
  if(shape == TRI) {
    /* RefTriRegDesc.C: 3 = mid(0,1), 4 = mid(1,2), 5 = mid(2,0) */
    const int e[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    const int ch[4][3] = {{0, 3, 5}, {1, 4, 3}, {2, 5, 4}, {3, 4, 5}};
    t.nref = 6; t.nchild = 4;
    for(int v = 0; v < 3; ++v) set_span(&t, v, 1, &v);
    for(int k = 0; k < 3; ++k) set_span(&t, 3 + k, 2, e[k]);
    for(int c = 0; c < 4; ++c) memcpy(t.child[c], ch[c], sizeof(ch[c]));
  } else if(shape == QUAD) {
    /* RefQuadRegDesc.C */
    const int e[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    const int all[4] = {0, 1, 2, 3};
    const int ch[4][4] = {{0, 4, 8, 7}, {1, 5, 8, 4}, {2, 6, 8, 5}, {3, 7, 8, 6}};
    t.nref = 9; t.nchild = 4;
    for(int v = 0; v < 4; ++v) set_span(&t, v, 1, &v);
    for(int k = 0; k < 4; ++k) set_span(&t, 4 + k, 2, e[k]);
    set_span(&t, 8, 4, all);
    for(int c = 0; c < 4; ++c) memcpy(t.child[c], ch[c], sizeof(ch[c]));
  } else if(shape == TET) {
    /* RefTetraRegDesc.C:15-16 and its edge-vertex table */
    const int e[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    const int ch[8][4] = {{0, 4, 6, 7}, {4, 1, 5, 8}, {6, 5, 2, 9}, {7, 8, 9, 3},
                          {4, 6, 7, 8}, {5, 6, 4, 8}, {6, 7, 8, 9}, {8, 5, 6, 9}};
    t.nref = 10; t.nchild = 8;
    for(int v = 0; v < 4; ++v) set_span(&t, v, 1, &v);
    for(int k = 0; k < 6; ++k) set_span(&t, 4 + k, 2, e[k]);
    for(int c = 0; c < 8; ++c) memcpy(t.child[c], ch[c], sizeof(ch[c]));
  } else {
    /* RefHexaRegDesc.C:17-23 and its edge/face vertex tables */
    static const int spans[27][9] = {
        {1, 0}, {1, 1}, {1, 2}, {1, 3}, {1, 4}, {1, 5}, {1, 6}, {1, 7},
        {2, 0, 1}, {2, 1, 2}, {2, 2, 3}, {2, 3, 0}, {4, 0, 1, 2, 3},
        {2, 1, 5}, {2, 5, 4}, {2, 4, 0}, {4, 0, 1, 5, 4},
        {2, 2, 6}, {2, 6, 5}, {4, 1, 2, 6, 5},
        {2, 3, 7}, {2, 7, 6}, {4, 2, 3, 7, 6},
        {2, 4, 7}, {4, 3, 0, 4, 7}, {4, 4, 5, 6, 7},
        {8, 0, 1, 2, 3, 4, 5, 6, 7}};
    const int ch[8][8] = {{0, 8, 12, 11, 15, 16, 26, 24}, {1, 9, 12, 8, 13, 19, 26, 16},
                          {2, 10, 12, 9, 17, 22, 26, 19}, {3, 11, 12, 10, 20, 24, 26, 22},
                          {4, 23, 25, 14, 15, 24, 26, 16}, {5, 14, 25, 18, 13, 16, 26, 19},
                          {6, 18, 25, 21, 17, 19, 26, 22}, {7, 21, 25, 23, 20, 22, 26, 24}};
    t.nref = 27; t.nchild = 8;
    memcpy(t.span, spans, sizeof(spans));
    for(int c = 0; c < 8; ++c) memcpy(t.child[c], ch[c], sizeof(ch[c]));
  }
  return t;
}

static void refine(const omesh* cm, omesh* fm) {
  reftab t = make_reftab(cm->shape);
  const int nv = NV[cm->shape];
  memset(fm, 0, sizeof(*fm));
  fm->shape = cm->shape;
  fm->ncells = cm->ncells * t.nchild;
  fm->cv = malloc(sizeof(int) * fm->ncells * nv);
  fm->parent = malloc(sizeof(int) * fm->ncells);
  fm->cap = cm->nverts * 8 + 64;
  fm->X = malloc(sizeof(double) * 3 * fm->cap);
  vdict d;
  d.cap = 1;
  while(d.cap < 4 * (cm->nverts * 8 + 64)) d.cap <<= 1;
  d.slot = malloc(sizeof(int) * d.cap);
  d.key = malloc(sizeof(int[4]) * d.cap);
  for(int i = 0; i < d.cap; ++i) d.slot[i] = -1;
  for(int c = 0; c < cm->ncells; ++c) {
    int gid[27];
    for(int r = 0; r < t.nref; ++r) {
      double x[3] = {0, 0, 0};
      const int n = t.span[r][0];
      for(int i = 0; i < n; ++i)
        for(int k = 0; k < 3; ++k) x[k] += cm->X[3 * cm->cv[c * nv + t.span[r][1 + i]] + k];
      for(int k = 0; k < 3; ++k) x[k] /= n;
      int key[4] = {-1, -1, -1, -1};
      if(n > 4) {
        key[0] = -2 - c; /* cell-interior vertex: never shared */
      } else {
        for(int i = 0; i < n; ++i) key[i] = cm->cv[c * nv + t.span[r][1 + i]];
        for(int a = 1; a < n; ++a) /* sort the spanning vertex ids */
          for(int b = a; b > 0 && key[b - 1] > key[b]; --b) { int tmp = key[b]; key[b] = key[b - 1]; key[b - 1] = tmp; }
      }
      gid[r] = vertex_id(fm, &d, key, x);
    }
    for(int ch = 0; ch < t.nchild; ++ch) {
      const int fc = c * t.nchild + ch;
      for(int a = 0; a < nv; ++a) fm->cv[fc * nv + a] = gid[t.child[ch][a]];
      fm->parent[fc] = c;
    }
  }
  free(d.slot);
  free(d.key);
}

/* ----------------------------------------------------------------------- */
/* adaptive refinement with hanging vertices (2D)                           */
/* ----------------------------------------------------------------------- */
/* TDomain::RefineByIndicator (src/Geometry/Domain.C:1094-1207,
   RefinementStrategy.C:260-287,
   GridCell.C:348-385): cells with corner values of the indicator on both
   sides of zero (|value| <= 1e-10 counts for both) are refined regularly,
   then cells next to a neighbour refined two levels deeper are refined
   until the mesh is 1-irregular. The leaves stay in depth-first tree order
   (a refined cell is replaced by its children in child order). Edge
   midpoints are kept in `mids` across steps, keyed by the sorted ids of the
   edge's end vertices, so that a cell refined later reuses the midpoint its
   neighbour created: a midpoint present on a leaf's edge is a hanging
   vertex of that leaf. */
typedef struct {
  double (*fn)(const double* x, void* ctx); /* indicator at a vertex */
  void* ctx;
} oindicator;

static int edge_mid(const vdict* mids, int u, int v) {
  int key[4] = {u < v ? u : v, u < v ? v : u, -1, -1};
  return vertex_find(mids, key);
}

static void refine_marked(const omesh* cm, const char* mark, omesh* fm, vdict* mids, int* center_count) {
  reftab t = make_reftab(cm->shape);
  const int nv = NV[cm->shape];
  int nref = 0;
  for(int c = 0; c < cm->ncells; ++c) nref += mark[c] != 0;
  memset(fm, 0, sizeof(*fm));
  fm->shape = cm->shape;
  fm->ncells = cm->ncells + nref * (t.nchild - 1);
  fm->cv = malloc(sizeof(int) * fm->ncells * nv);
  fm->parent = malloc(sizeof(int) * fm->ncells);
  fm->cap = cm->nverts + nref * 8 + 64;
  fm->X = malloc(sizeof(double) * 3 * fm->cap);
  memcpy(fm->X, cm->X, sizeof(double) * 3 * cm->nverts);
  fm->nverts = cm->nverts;
  int fc = 0;
  for(int c = 0; c < cm->ncells; ++c) {
    if(!mark[c]) {
      memcpy(&fm->cv[fc * nv], &cm->cv[c * nv], sizeof(int) * nv);
      fm->parent[fc++] = c;
      continue;
    }
    int gid[27];
    for(int r = 0; r < t.nref; ++r) {
      const int n = t.span[r][0];
      if(n == 1) { gid[r] = cm->cv[c * nv + t.span[r][1]]; continue; }
      double x[3] = {0, 0, 0};
      for(int i = 0; i < n; ++i)
        for(int k = 0; k < 3; ++k) x[k] += cm->X[3 * cm->cv[c * nv + t.span[r][1 + i]] + k];
      for(int k = 0; k < 3; ++k) x[k] /= n;
      int key[4] = {-1, -1, -1, -1};
      if(n == 2) {
        const int u = cm->cv[c * nv + t.span[r][1]], v = cm->cv[c * nv + t.span[r][2]];
        key[0] = u < v ? u : v;
        key[1] = u < v ? v : u;
      } else {
        key[0] = -2 - (*center_count)++; /* cell-interior vertex: never shared */
      }
      gid[r] = vertex_id(fm, mids, key, x);
    }
    for(int ch = 0; ch < t.nchild; ++ch) {
      for(int a = 0; a < nv; ++a) fm->cv[fc * nv + a] = gid[t.child[ch][a]];
      fm->parent[fc++] = c;
    }
  }
}

/* the leaves after `steps` indicator refinements of `base` (consumed);
   mids keeps the edge midpoints */
static int straddles(const oindicator* ind, const double* const* pts, int n) {
  int inner = 0, outer = 0;
  for(int a = 0; a < n; ++a) {
    const double val = ind->fn(pts[a], ind->ctx);
    if(val <= 1e-10) inner++;
    if(val >= -1e-10) outer++;
  }
  return inner > 0 && outer > 0;
}

/* refines the marked leaves (consumes mark), then every leaf whose edge
   (u, v) has a midpoint m with (u, m) or (m, v) split again, until the mesh
   is 1-irregular */
static void refine_1irregular(omesh* cur, char* mark, vdict* mids, int* centers) {
  const int nv = NV[cur->shape];
  for(;;) {
    omesh next;
    refine_marked(cur, mark, &next, mids, centers);
    mesh_free(cur);
    free(mark);
    *cur = next;
    mark = calloc(cur->ncells, 1);
    int any = 0;
    for(int c = 0; c < cur->ncells; ++c)
      for(int j = 0; j < nv; ++j) {
        const int u = cur->cv[c * nv + j], v = cur->cv[c * nv + (j + 1) % nv];
        const int mm = edge_mid(mids, u, v);
        if(mm >= 0 && (edge_mid(mids, u, mm) >= 0 || edge_mid(mids, mm, v) >= 0)) {
          mark[c] = 1;
          any = 1;
        }
      }
    if(!any || orc_fault == 8) break;
  }
  free(mark);
}

/* children of a triangle bisected on edge j (vertices v_j, v_j+1, midpoint
   m): {v_j, m, v_j+2} and {v_j+1, v_j+2, m} (RefTriBis0Desc.C:21-22,
   RefTriBis1Desc.C:22-23, RefTriBis2Desc.C:22-23) */
static void bisect_children(const int* v, int j, int m, int ch[2][3]) {
  ch[0][0] = v[j]; ch[0][1] = m; ch[0][2] = v[(j + 2) % 3];
  ch[1][0] = v[(j + 1) % 3]; ch[1][1] = v[(j + 2) % 3]; ch[1][2] = m;
}

/* `steps` indicator refinements of `base` (consumed); mids keeps the edge
   midpoints. conforming (triangles): TDomain::MakeConfClosure
   (src/Geometry/Domain.C:1219-1360, GridCell.C:525-575): before a step the
   closure bisections are removed (a bisected cell is marked when one of its
   two halves is); after the 1-irregular refinement, every triangle with two
   or three refined edges (edges whose neighbour across is refined) is
   refined regularly, until none is left, and every triangle with one is
   bisected on it (the bisections are the leaves' last level, rebuilt each
   step). */
static void adapt_mesh(omesh* base, const oindicator* ind, int steps, omesh* out, vdict* mids, int conforming) {
  mids->cap = 1 << 20;
  mids->slot = malloc(sizeof(int) * mids->cap);
  mids->key = malloc(sizeof(int[4]) * mids->cap);
  for(int i = 0; i < mids->cap; ++i) mids->slot[i] = -1;
  int centers = 0;
  omesh cur = *base;
  const int nv = NV[cur.shape];
  int* bis = NULL; /* conforming: bisected edge of each leaf of cur, or -1 */
  for(int step = 0; step < steps; ++step) {
    char* mark = calloc(cur.ncells, 1);
    for(int c = 0; c < cur.ncells; ++c) {
      const int* v = &cur.cv[c * nv];
      if(bis && bis[c] >= 0) {
        const int j = bis[c];
        const int m = edge_mid(mids, v[j], v[(j + 1) % 3]);
        int ch[2][3];
        bisect_children(v, j, m, ch);
        for(int h = 0; h < 2; ++h) {
          const double* pts[3] = {&cur.X[3 * ch[h][0]], &cur.X[3 * ch[h][1]], &cur.X[3 * ch[h][2]]};
          if(straddles(ind, pts, 3)) mark[c] = 1;
        }
      } else {
        const double* pts[8];
        for(int a = 0; a < nv; ++a) pts[a] = &cur.X[3 * v[a]];
        mark[c] = straddles(ind, pts, nv);
      }
    }
    refine_1irregular(&cur, mark, mids, &centers);
    free(bis);
    bis = NULL;
    if(!conforming) continue;
    /* closure: regular refinement of triangles with two or more refined edges */
    for(;;) {
      mark = calloc(cur.ncells, 1);
      int any = 0;
      for(int c = 0; c < cur.ncells; ++c) {
        int n = 0;
        for(int j = 0; j < 3; ++j) n += edge_mid(mids, cur.cv[c * 3 + j], cur.cv[c * 3 + (j + 1) % 3]) >= 0;
        if(n >= (orc_fault == 11 ? 3 : 2)) { mark[c] = 1; any = 1; }
      }
      if(!any) { free(mark); break; }
      refine_1irregular(&cur, mark, mids, &centers);
    }
    bis = malloc(sizeof(int) * cur.ncells);
    for(int c = 0; c < cur.ncells; ++c) {
      bis[c] = -1;
      for(int j = 0; j < 3; ++j)
        if(edge_mid(mids, cur.cv[c * 3 + j], cur.cv[c * 3 + (j + 1) % 3]) >= 0) bis[c] = j;
    }
  }
  if(bis) {
    /* the leaves with the bisected cells replaced by their halves */
    int n = 0;
    for(int c = 0; c < cur.ncells; ++c) n += bis[c] >= 0 ? 2 : 1;
    omesh fm = cur;
    fm.ncells = n;
    fm.cv = malloc(sizeof(int) * 3 * n);
    fm.parent = malloc(sizeof(int) * n);
    int f = 0;
    for(int c = 0; c < cur.ncells; ++c) {
      const int* v = &cur.cv[c * 3];
      if(bis[c] < 0) {
        memcpy(&fm.cv[3 * f], v, sizeof(int) * 3);
        fm.parent[f++] = c;
        continue;
      }
      int ch[2][3];
      bisect_children(v, bis[c], edge_mid(mids, v[bis[c]], v[(bis[c] + 1) % 3]), ch);
      for(int h = 0; h < 2; ++h) {
        memcpy(&fm.cv[3 * f], ch[h], sizeof(int) * 3);
        fm.parent[f++] = c;
      }
    }
    free(cur.cv);
    free(cur.parent);
    cur = fm;
    free(bis);
  }
  *out = cur;
}

/* boundary component of a joint of the unit box from its vertices
   (Default_UnitSquare / Default_UnitCube parameter files) */
static int box_comp(int dim, const double* xs, int n) {
  for(int plane = 0; plane < (dim == 2 ? 4 : 6); ++plane) {
    int ok = 1;
    for(int i = 0; i < n && ok; ++i) {
      const double* x = xs + 3 * i;
      if(dim == 2) {
        const double v[4] = {x[1], x[0] - 1, x[1] - 1, x[0]};
        ok = v[plane] == 0.0;
      } else {
        const double v[6] = {x[2], x[0], x[1], x[2] - 1, x[0] - 1, x[1] - 1};
        ok = v[plane] == 0.0;
      }
    }
    if(ok) return plane;
  }
  return -1;
}

/* ----------------------------------------------------------------------- */
/* elements: node sets (FE descriptor order) and Lagrange bases             */
/* ----------------------------------------------------------------------- */
typedef struct {
  int shape, order, nl, dim;
  double node[MAXNL][3];
  int simplex_alpha[MAXNL][4]; /* barycentric multi-index (simplices) */
  int tensor_idx[MAXNL][3];    /* 1D node indices (tensor elements) */
} oelem;

static oelem make_elem(int shape, int order) {
  oelem e;
  memset(&e, 0, sizeof(e));
  e.shape = shape;
  e.order = order;
  e.dim = DIM[shape];
  if(shape == QUAD || shape == HEX) {
    const int n1 = order + 1, nz = e.dim == 3 ? n1 : 1;
    int i = 0;
    for(int c = 0; c < nz; ++c)
      for(int b = 0; b < n1; ++b)
        for(int a = 0; a < n1; ++a, ++i) {
          e.tensor_idx[i][0] = a;
          e.tensor_idx[i][1] = b;
          e.tensor_idx[i][2] = c;
          e.node[i][0] = -1.0 + 2.0 * a / order;
          e.node[i][1] = -1.0 + 2.0 * b / order;
          e.node[i][2] = e.dim == 3 ? -1.0 + 2.0 * c / order : 0.0;
        }
    e.nl = i;
  } else {
    /* P1: vertices; P2 (FE_C_T_P2_2D.h / FE_C_T_P2_3D.h layout): nodes in
       lexicographic order of (zeta, eta, xi) on the grid of step 1/2 */
    int i = 0;
    const int p = order;
    const int nz = e.dim == 3 ? p : 0;
    for(int c = 0; c <= nz; ++c)
      for(int b = 0; b <= p - c; ++b)
        for(int a = 0; a <= p - b - c; ++a, ++i) {
          e.node[i][0] = (double)a / p;
          e.node[i][1] = (double)b / p;
          e.node[i][2] = (double)c / p;
          e.simplex_alpha[i][1] = a;
          e.simplex_alpha[i][2] = b;
          e.simplex_alpha[i][3] = c;
          e.simplex_alpha[i][0] = p - a - b - c;
        }
    e.nl = i;
  }
  return e;
}

/* 1D Lagrange polynomial through equidistant nodes of [-1,1]: value and
   first two derivatives, by the product formula */
static void lag1d(int order, int a, double x, double* v, double* d1, double* d2) {
  double nodes[3];
  for(int k = 0; k <= order; ++k) nodes[k] = -1.0 + 2.0 * k / order;
  double val = 1, der = 0, der2 = 0;
  /* expand product incrementally: (f g)' = f' g + f g', (f g)'' = f'' g + 2 f' g' + f g'' */
  for(int k = 0; k <= order; ++k) {
    if(k == a) continue;
    const double den = nodes[a] - nodes[k];
    const double g = (x - nodes[k]) / den, gd = 1.0 / den;
    der2 = der2 * g + 2 * der * gd;
    der = der * g + val * gd;
    val = val * g;
  }
  *v = val;
  *d1 = der;
  *d2 = der2;
}

/* Silvester's factor P_a(l) = prod_{k<a} (p l - k)/(k+1) with derivatives */
static void silv(int p, int a, double l, double* v, double* d1, double* d2) {
  double val = 1, der = 0, der2 = 0;
  for(int k = 0; k < a; ++k) {
    const double g = (p * l - k) / (k + 1), gd = (double)p / (k + 1);
    der2 = der2 * g + 2 * der * gd;
    der = der * g + val * gd;
    val = val * g;
  }
  *v = val;
  *d1 = der;
  *d2 = der2;
}

/* out[c*nl+i], c over D0, first derivatives, then second derivatives in
   multi-index order (2D: xx, xy, yy; 3D: xx, xy, xz, yy, yz, zz) */
static void elem_eval(const oelem* e, const double* p, double* out) {
  const int nl = e->nl, d = e->dim;
  const int pr2[3][2] = {{0, 0}, {0, 1}, {1, 1}};
  const int pr3[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  const int n2 = d == 2 ? 3 : 6;
  for(int i = 0; i < nl; ++i) {
    double grad[3] = {0, 0, 0}, hess[6] = {0, 0, 0, 0, 0, 0}, val;
    if(e->shape == QUAD || e->shape == HEX) {
      double v[3], a1[3], a2[3];
      for(int k = 0; k < d; ++k) lag1d(e->order, e->tensor_idx[i][k], p[k], &v[k], &a1[k], &a2[k]);
      val = 1;
      for(int k = 0; k < d; ++k) val *= v[k];
      for(int k = 0; k < d; ++k) {
        double g = a1[k];
        for(int m = 0; m < d; ++m)
          if(m != k) g *= v[m];
        grad[k] = g;
      }
      for(int t = 0; t < n2; ++t) {
        const int k = d == 2 ? pr2[t][0] : pr3[t][0], l = d == 2 ? pr2[t][1] : pr3[t][1];
        double h = 1;
        for(int m = 0; m < d; ++m) {
          if(k == l && m == k) h *= a2[m];
          else if(m == k || m == l) h *= a1[m];
          else h *= v[m];
        }
        hess[t] = h;
      }
    } else {
      double lam[4], f[4], f1[4], f2[4];
      lam[0] = 1;
      for(int k = 0; k < d; ++k) { lam[0] -= p[k]; lam[k + 1] = p[k]; }
      for(int j = 0; j <= d; ++j) silv(e->order, e->simplex_alpha[i][j], lam[j], &f[j], &f1[j], &f2[j]);
      /* d lam_j / d xi_k */
      double g[4][3] = {{0}};
      for(int k = 0; k < d; ++k) { g[0][k] = -1; g[k + 1][k] = 1; }
      val = 1;
      for(int j = 0; j <= d; ++j) val *= f[j];
      double dl[4], ddl[4][4];
      for(int j = 0; j <= d; ++j) {
        double t = f1[j];
        for(int m = 0; m <= d; ++m)
          if(m != j) t *= f[m];
        dl[j] = t;
        for(int jj = 0; jj <= d; ++jj) {
          double h = 1;
          for(int m = 0; m <= d; ++m) {
            if(j == jj && m == j) h *= f2[m];
            else if(m == j || m == jj) h *= f1[m];
            else h *= f[m];
          }
          ddl[j][jj] = h;
        }
      }
      for(int k = 0; k < d; ++k) {
        double s = 0;
        for(int j = 0; j <= d; ++j) s += g[j][k] * dl[j];
        grad[k] = s;
      }
      for(int t = 0; t < n2; ++t) {
        const int k = d == 2 ? pr2[t][0] : pr3[t][0], l = d == 2 ? pr2[t][1] : pr3[t][1];
        double s = 0;
        for(int j = 0; j <= d; ++j)
          for(int jj = 0; jj <= d; ++jj) s += g[j][k] * g[jj][l] * ddl[j][jj];
        hess[t] = s;
      }
    }
    out[i] = val;
    for(int k = 0; k < d; ++k) out[(1 + k) * nl + i] = grad[k];
    for(int t = 0; t < n2; ++t) out[(1 + d + t) * nl + i] = hess[t];
  }
}

/* ----------------------------------------------------------------------- */
/* quadrature (QuadFormula.C data, degree map QuadratureFormulaDatabase.C)  */
/* ----------------------------------------------------------------------- */
typedef struct {
  int n;
  double w[MAXQ], x[MAXQ][3];
} oquad;

static oquad make_quad(int shape, int degree) {
  oquad q;
  memset(&q, 0, sizeof(q));
  if(shape == TRI) {
    if(degree <= 5) {
      const double a = 0.797426985353087322398025276169754, b = 0.101286507323456338800987361915123;
      const double c = 0.059715871789769820459117580973106, dd = 0.470142064105115089770441209513447;
      const double wa = 0.0629695902724135762978419727500906, wc = 0.0661970763942530903688246939165759;
      const double W[7] = {0.1125, wa, wa, wa, wc, wc, wc};
      const double X[7] = {0.333333333333333333333333333333333, a, b, b, c, dd, dd};
      const double Y[7] = {0.333333333333333333333333333333333, b, a, b, dd, c, dd};
      q.n = 7;
      for(int i = 0; i < 7; ++i) { q.w[i] = W[i]; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; }
    } else {
      const double w0 = 0.0485678981413994169096209912536443, w1 = 0.0156673501135695352684274156436046,
                   w2 = 0.0389137705023871396583696781497019, w3 = 0.0398238694636051265164458871320226,
                   w4 = 0.0127888378293490156308393992794999, w5 = 0.0216417696886446886446886446886446;
      const double a1 = 0.489682519198737627783706924836192, b1 = 0.020634961602524744432586150327616;
      const double a2 = 0.437089591492936637269930364435354, b2 = 0.125820817014126725460139271129292;
      const double a3 = 0.188203535619032730240961280467335, b3 = 0.623592928761934539518077439065330;
      const double a4 = 0.0447295133944527098651065899662763, b4 = 0.910540973211094580269786820067447;
      const double p = 0.741198598784498020690079873523423, r = 0.0368384120547362836348175987833851,
                   s = 0.221962989160765695675102527693192, th = 0.333333333333333333333333333333333;
      const double W[19] = {w0, w1, w1, w1, w2, w2, w2, w3, w3, w3, w4, w4, w4, w5, w5, w5, w5, w5, w5};
      const double X[19] = {th, a1, a1, b1, a2, a2, b2, a3, a3, b3, a4, a4, b4, p, r, p, s, r, s};
      const double Y[19] = {th, a1, b1, a1, a2, b2, a2, a3, b3, a3, a4, b4, a4, r, p, s, p, s, r};
      q.n = 19;
      for(int i = 0; i < 19; ++i) { q.w[i] = W[i]; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; }
    }
  } else if(shape == QUAD) {
    if(degree <= 2) {
      const double g = 0.57735026918962576450914878050195746;
      const double X[4] = {-g, g, -g, g}, Y[4] = {-g, -g, g, g};
      q.n = 4;
      for(int i = 0; i < 4; ++i) { q.w[i] = 1; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; }
    } else if(degree <= 5) {
      const double g = 0.77459666924148337703585307995647992;
      const double wa = 0.30864197530864197530864197530864198, wb = 0.49382716049382716049382716049382717,
                   wc = 0.79012345679012345679012345679012346;
      const double W[9] = {wa, wb, wa, wb, wc, wb, wa, wb, wa};
      const double X[9] = {-g, 0., g, -g, 0., g, -g, 0., g}, Y[9] = {-g, -g, -g, 0., 0., 0., g, g, g};
      q.n = 9;
      for(int i = 0; i < 9; ++i) { q.w[i] = W[i]; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; }
    } else {
      const double a = 0.86113631159405257522394648889280951, b = 0.33998104358485626480266575910324469;
      const double wa = 0.12100299328560200552121209737014755, wb = 0.22685185185185185185185185185185185,
                   wc = 0.42529330301069429077508419892614875;
      const double W[16] = {wa, wb, wb, wa, wb, wc, wc, wb, wb, wc, wc, wb, wa, wb, wb, wa};
      const double x1[4] = {-a, -b, b, a};
      q.n = 16;
      for(int i = 0; i < 16; ++i) { q.w[i] = W[i]; q.x[i][0] = x1[i % 4]; q.x[i][1] = x1[i / 4]; }
    }
  } else if(shape == TET) {
    if(degree <= 2) {
      const double a = 0.138196601125010515179541316563, b = 0.585410196624968454461376050311;
      const double X[4] = {a, a, a, b}, Y[4] = {a, a, b, a}, Z[4] = {a, b, a, a};
      q.n = 4;
      for(int i = 0; i < 4; ++i) { q.w[i] = 1. / 24.; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; q.x[i][2] = Z[i]; }
    } else if(degree <= 4) {
      const double w0 = -.0131555555555555555555555555556, w1 = .00762222222222222222222222222222,
                   w2 = .0248888888888888888888888888889;
      const double s = .0714285714285714285714285714286, t = .785714285714285714285714285714;
      const double u = .399403576166799204996102147462, v = .100596423833200795003897852538;
      const double W[11] = {w0, w1, w1, w1, w1, w2, w2, w2, w2, w2, w2};
      const double X[11] = {.25, s, t, s, s, u, u, v, v, v, u};
      const double Y[11] = {.25, s, s, t, s, u, v, u, v, u, v};
      const double Z[11] = {.25, s, s, s, t, v, u, u, u, v, v};
      q.n = 11;
      for(int i = 0; i < 11; ++i) { q.w[i] = W[i]; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; q.x[i][2] = Z[i]; }
    } else {
      const double w0 = 0.602678571428571428571428571428571e-2, w1 = 0.302836780970891758063769725577305e-1,
                   w2 = 0.116452490860289694108936091443380e-1, w3 = 0.109491415613864593456430191124068e-1;
      const double th = 0.333333333333333333333333333333333;
      const double a = 0.909090909090909090909090909090909e-1, b = 0.727272727272727272727272727272727;
      const double c = 0.665501535736642982398804642263025e-1, d = 0.433449846426335701760119535773697;
      const double W[15] = {w0, w0, w0, w0, w1, w2, w2, w2, w2, w3, w3, w3, w3, w3, w3};
      const double X[15] = {th, th, 0, th, 0.25, a, a, b, a, c, d, d, c, c, d};
      const double Y[15] = {th, 0, th, th, 0.25, a, b, a, a, d, c, d, c, d, c};
      const double Z[15] = {0, th, th, th, 0.25, b, 0.90909090909090909090909090909091e-1,
                            0.909090909090909090909090909090911e-1, 0.909090909090909090909090909090911e-1,
                            0.433449846426335701760119535773699, 0.433449846426335701760119535773699,
                            0.66550153573664298239880464226304e-1, 0.433449846426335701760119535773699,
                            0.66550153573664298239880464226304e-1, 0.665501535736642982398804642263035e-1};
      q.n = 15;
      for(int i = 0; i < 15; ++i) { q.w[i] = W[i]; q.x[i][0] = X[i]; q.x[i][1] = Y[i]; q.x[i][2] = Z[i]; }
    }
  } else {
    double w1[4], x1[4];
    int n1;
    if(degree <= 2) {
      const double g = 0.5773502691896257645091489;
      n1 = 2; w1[0] = w1[1] = 1.0; x1[0] = -g; x1[1] = g;
    } else if(degree <= 5) {
      n1 = 3;
      w1[0] = 5. / 9.; w1[1] = 8. / 9.; w1[2] = 5. / 9.;
      x1[0] = -0.7745966692414833770358530; x1[1] = 0; x1[2] = 0.7745966692414833770358530;
    } else {
      n1 = 4;
      w1[0] = 0.652145154862546142626936051; w1[1] = 0.347854845137453857373063949;
      w1[2] = 0.347854845137453857373063949; w1[3] = 0.652145154862546142626936051;
      x1[0] = 0.33998104358485626480266576; x1[1] = 0.86113631159405257522394649;
      x1[2] = -0.86113631159405257522394649; x1[3] = -0.33998104358485626480266576;
    }
    q.n = 0;
    for(int i = 0; i < n1; ++i)
      for(int j = 0; j < n1; ++j)
        for(int k = 0; k < n1; ++k) {
          q.w[q.n] = w1[i] * w1[j] * w1[k];
          q.x[q.n][0] = x1[k];
          q.x[q.n][1] = x1[j];
          q.x[q.n][2] = x1[i];
          q.n++;
        }
  }
  return q;
}

/* ----------------------------------------------------------------------- */
/* examples (problem data)                                                  */
/* ----------------------------------------------------------------------- */
typedef struct {
  int kind;
  double eps, b[3], c, f0;
  double t; /* time of the time-dependent examples 7, 8 */
} oexample;

/* coeffs: eps, b1, b2, b3, c, f */
static void ex_coeffs(const oexample* ex, double x, double y, double z, double* o) {
  const double p = M_PI;
  switch(ex->kind) {
    case 0:
      o[0] = ex->eps; o[1] = o[2] = o[3] = 0; o[4] = 0;
      o[5] = -o[0] * (-2 * p * p * sin(p * x) * sin(p * y));
      break;
    case 1: {
      double ux = p * cos(p * x) * sin(p * y), uy = p * sin(p * x) * cos(p * y);
      o[0] = ex->eps; o[1] = ex->b[0]; o[2] = ex->b[1]; o[3] = 0; o[4] = 0;
      o[5] = -o[0] * (-2 * p * p * sin(p * x) * sin(p * y)) + o[1] * ux + o[2] * uy;
      break;
    }
    case 2: {
      double sx = sin(p * x), sy = sin(p * y), sz = sin(p * z);
      o[0] = ex->eps; o[1] = ex->b[0]; o[2] = ex->b[1]; o[3] = ex->b[2]; o[4] = 0;
      o[5] = 3 * p * p * ex->eps * (sx * sy * sz) + ex->b[0] * (p * cos(p * x) * sy * sz) +
             ex->b[1] * (p * sx * cos(p * y) * sz) + ex->b[2] * (p * sx * sy * cos(p * z));
      break;
    }
    case 3: o[0] = 1; o[1] = 1; o[2] = 1; o[3] = 1; o[4] = 1; o[5] = 1; break;
    case 4: o[0] = 1; o[1] = 1; o[2] = 2; o[3] = 3; o[4] = 0; o[5] = 20; break;
    case 5:
      o[0] = 1; o[1] = 4; o[2] = 3; o[3] = 2; o[4] = 1;
      o[5] = -1 * 2 + 4 * 2 * x + 3 * z + 2 * y + 1 * (x * x + y * z);
      break;
    case 7: { /* include/Examples/TCD_3D/linear_space_time.h */
      const double t = ex->t;
      o[0] = 1; o[1] = 1; o[2] = 2; o[3] = 3; o[4] = 0;
      o[5] = 3 * y + 4 * z + 2 + 6 * t + 12 * t;
      break;
    }
    case 8: { /* include/Examples/TCD_3D/quadratic_space_time.h */
      const double t = ex->t, eps = 1;
      o[0] = eps; o[1] = 4; o[2] = 3; o[3] = 2; o[4] = 1;
      o[5] = 2 * t * (x * x + y * z) - eps * 2 * t * t + 4 * 2 * x * t * t + 3 * z * t * t + 2 * y * t * t +
             1 * t * t * (x * x + y * z);
      break;
    }
    case 11: { /* include/Examples/CD_3D/smooth_solution_3d.h: f = -eps lap u + b.grad u + c u */
      double e[4];
      const double lap = -2 * y * z * (1 - y) * (1 - z) - 2 * x * z * (1 - x) * (1 - z) - 2 * x * y * (1 - x) * (1 - y);
      o[0] = ex->eps; o[1] = 1; o[2] = 1; o[3] = 1; o[4] = 1;
      e[0] = x * y * z * (1 - x) * (1 - y) * (1 - z);
      e[1] = y * z * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - y) * (1 - z);
      e[2] = x * z * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - x) * (1 - z);
      e[3] = x * y * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - x) * (1 - y);
      o[5] = -o[0] * lap;
      o[5] += o[1] * e[1] + o[2] * e[2] + o[3] * e[3];
      o[5] += o[4] * e[0];
      break;
    }
    case 10: /* include/Examples/CD_2D/SharpBoundaryLayer.h */
      o[0] = 0.001; o[1] = cos(M_PI / 18); o[2] = sin(M_PI / 18); o[3] = 0; o[4] = 0; o[5] = 0;
      break;
    case 9: { /* include/Examples/TCD_2D/exp.h */
      const double q = 2 * p, et = exp(sin(q * ex->t)), s = sin(q * x) * sin(q * y);
      o[0] = ex->eps; o[1] = 1; o[2] = -1; o[3] = 0; o[4] = 1;
      o[5] = et * q * cos(q * ex->t) * s + 2 * ex->eps * q * q * et * s + et * q * cos(q * x) * sin(q * y) -
             et * q * sin(q * x) * cos(q * y) + et * s;
      break;
    }
    default:
      o[0] = ex->eps; o[1] = ex->b[0]; o[2] = ex->b[1]; o[3] = ex->b[2]; o[4] = ex->c; o[5] = ex->f0;
  }
}

/* exact: u, ux, uy, uz; returns 0 if unknown */
static int ex_exact(const oexample* ex, double x, double y, double z, double* o) {
  const double p = M_PI;
  switch(ex->kind) {
    case 0:
    case 1:
      o[0] = sin(p * x) * sin(p * y); o[1] = p * cos(p * x) * sin(p * y);
      o[2] = p * sin(p * x) * cos(p * y); o[3] = 0;
      return 1;
    case 2:
      o[0] = sin(p * x) * sin(p * y) * sin(p * z); o[1] = p * cos(p * x) * sin(p * y) * sin(p * z);
      o[2] = p * sin(p * x) * cos(p * y) * sin(p * z); o[3] = p * sin(p * x) * sin(p * y) * cos(p * z);
      return 1;
    case 3: o[0] = 1; o[1] = o[2] = o[3] = 0; return 1;
    case 4: o[0] = 1 + 2 * x + 3 * y + 4 * z; o[1] = 2; o[2] = 3; o[3] = 4; return 1;
    case 5: o[0] = x * x + y * z; o[1] = 2 * x; o[2] = z; o[3] = y; return 1;
    case 7:
      o[0] = 1 + 2 * x + 3 * ex->t * y + 4 * ex->t * z; o[1] = 2; o[2] = 3 * ex->t; o[3] = 4 * ex->t;
      return 1;
    case 8: {
      const double t = ex->t;
      o[0] = t * t * (x * x + y * z); o[1] = 2 * x * t * t; o[2] = z * t * t; o[3] = y * t * t;
      return 1;
    }
    case 11:
      o[0] = x * y * z * (1 - x) * (1 - y) * (1 - z);
      o[1] = y * z * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - y) * (1 - z);
      o[2] = x * z * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - x) * (1 - z);
      o[3] = x * y * (1 - x) * (1 - y) * (1 - z) - x * y * z * (1 - x) * (1 - y);
      return 1;
    case 9: {
      const double q = 2 * p, et = exp(sin(q * ex->t));
      o[0] = et * sin(q * x) * sin(q * y); o[1] = et * q * cos(q * x) * sin(q * y);
      o[2] = et * q * sin(q * x) * cos(q * y); o[3] = 0;
      return 1;
    }
    default: o[0] = o[1] = o[2] = o[3] = 0; return 0;
  }
}

static int ex_bc(const oexample* ex, int comp) {
  return ((ex->kind == 0 && comp == 1) || (ex->kind == 10 && comp == 2)) ? 1 : 0;
}

static double ex_bvalue(const oexample* ex, int comp, double x, double y, double z) {
  double e[4];
  if(ex->kind == 9 || ex->kind == 11) return 0; /* exp.h, smooth_solution_3d.h BoundValue */
  /* SharpBoundaryLayer.h BoundValue: 1 on x = 0 for parameter < 0.5 (y > 0.5),
     0 elsewhere and on the Neumann side y = 1 */
  if(ex->kind == 10) return (comp != 2 && fabs(x) < 1e-12 && y > 0.5) ? 1.0 : 0.0;
  if(!ex_exact(ex, x, y, z, e)) return 0;
  if(ex->kind == 0 && comp == 1) return ex->eps * (1.0 * e[1] + 0.0 * e[2]);
  return e[0];
}

/* ----------------------------------------------------------------------- */
/* level: mesh + space + matrix                                             */
/* ----------------------------------------------------------------------- */
typedef struct {
  omesh m;
  int nl, ndof, nactive, ndir;
  int* dof; /* GlobalNumbers */
  int* rp;
  int* col;
  long nnz;
  double* val;
  double* rhs;
  double* sol;
  /* multigrid work */
  double *mg_sol, *mg_rhs, *mg_def, *mg_aux;
  double* lu; /* dense LU of the coarse/direct matrix */
  int* piv;
  double* qq;   /* per child number: coarse basis at the child's nodes */
  int qq_ready[8];
  double* diag; /* diagonal entries (Iteration_jacobi's copy) */
  int* nb_cell; /* face neighbours: nb_cell[c*NJ+j], -1 boundary, -2 interior
                   joint without a conforming neighbour (hanging vertex) */
  int* nb_joint;
  int* comp;    /* boundary component per joint, -1 interior */
  double *aval, *mval; /* time-dependent problems: stiffness and mass matrices */
  /* hanging DOFs nactive_nh .. nactive-1: per hanging DOF hang_ncpl[h]
     coupling DOFs hang_cpl[3h + k] with weights hang_w[3h + k], and the
     topological order of the list */
  int nhang, nactive_nh;
  int* hang_ncpl;
  int* hang_cpl;
  double* hang_w;
  int* hang_order;
} olevel;

typedef struct orc_problem {
  int nlev;
  olevel* lev; /* 0 = coarsest */
  oelem el;
  oquad qd;
  oexample ex;
  int supg;
  /* time-dependent problems (TimeConvectionDiffusion + TimeDiscretization) */
  int tcd, scheme, bdf_first, step;
  int afc; /* algebraic flux correction: Dirichlet rows assembled like the others */
  double tau, time;
  double *u_old, *u_old2, *f_old;
} orc_problem;

/* face neighbours by matching sorted vertex ids of joints */
typedef struct { int v[4]; int cell, joint; } facekey;
static int fk_cmp(const void* a, const void* b) {
  const facekey* x = a;
  const facekey* y = b;
  for(int k = 0; k < 4; ++k)
    if(x->v[k] != y->v[k]) return x->v[k] < y->v[k] ? -1 : 1;
  if(x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
  return x->joint - y->joint;
}

static void build_neighbours(olevel* L) {
  const omesh* m = &L->m;
  const int s = m->shape, nj = NJ[s], nv = NV[s], len = jv_len(s);
  const int n = m->ncells * nj;
  facekey* fk = malloc(sizeof(facekey) * n);
  for(int c = 0; c < m->ncells; ++c)
    for(int j = 0; j < nj; ++j) {
      facekey* f = &fk[c * nj + j];
      int v[4] = {-1, -1, -1, -1};
      for(int k = 0; k < len; ++k) v[k] = m->cv[c * nv + jv(s, j, k)];
      /* sort */
      for(int a = 0; a < len; ++a)
        for(int b = a + 1; b < len; ++b)
          if(v[b] < v[a]) { int t = v[a]; v[a] = v[b]; v[b] = t; }
      memcpy(f->v, v, sizeof(v));
      f->cell = c;
      f->joint = j;
    }
  qsort(fk, n, sizeof(facekey), fk_cmp);
  L->nb_cell = malloc(sizeof(int) * n);
  L->nb_joint = malloc(sizeof(int) * n);
  L->comp = malloc(sizeof(int) * n);
  for(int i = 0; i < n; ++i) { L->nb_cell[i] = -1; L->nb_joint[i] = -1; }
  for(int i = 0; i + 1 < n; ++i)
    if(memcmp(fk[i].v, fk[i + 1].v, sizeof(fk[i].v)) == 0) {
      L->nb_cell[fk[i].cell * nj + fk[i].joint] = fk[i + 1].cell;
      L->nb_joint[fk[i].cell * nj + fk[i].joint] = fk[i + 1].joint;
      L->nb_cell[fk[i + 1].cell * nj + fk[i + 1].joint] = fk[i].cell;
      L->nb_joint[fk[i + 1].cell * nj + fk[i + 1].joint] = fk[i].joint;
    }
  free(fk);
  for(int c = 0; c < m->ncells; ++c)
    for(int j = 0; j < nj; ++j) {
      L->comp[c * nj + j] = -1;
      if(L->nb_cell[c * nj + j] >= 0) continue;
      double xs[12];
      for(int k = 0; k < len; ++k) memcpy(&xs[3 * k], &m->X[3 * m->cv[c * nv + jv(s, j, k)]], 24);
      L->comp[c * nj + j] = box_comp(DIM[s], xs, len);
      if(L->comp[c * nj + j] < 0) L->nb_cell[c * nj + j] = -2;
    }
}

/* physical coordinates of a reference point (affine maps of the reference) */
typedef struct {
  double c0[3], C[3][3], det, K[3][3], hK;
} ogeo;

static void geo(const omesh* m, int c, ogeo* g) {
  const int s = m->shape, nv = NV[s];
  double X[8][3];
  for(int a = 0; a < nv; ++a) memcpy(X[a], &m->X[3 * m->cv[c * nv + a]], 24);
  memset(g, 0, sizeof(*g));
  if(s == TRI || s == QUAD) {
    for(int a = 0; a < 2; ++a) {
      if(s == TRI) {
        g->c0[a] = X[0][a]; g->C[a][0] = X[1][a] - X[0][a]; g->C[a][1] = X[2][a] - X[0][a];
      } else {
        g->c0[a] = (X[1][a] + X[3][a]) * 0.5;
        g->C[a][0] = (X[1][a] - X[0][a]) * 0.5;
        g->C[a][1] = (X[3][a] - X[0][a]) * 0.5;
      }
    }
    const double xc1 = g->C[0][0], xc2 = g->C[0][1], yc1 = g->C[1][0], yc2 = g->C[1][1];
    g->det = xc1 * yc2 - xc2 * yc1;
    g->K[0][0] = yc2; g->K[0][1] = -yc1; g->K[1][0] = -xc2; g->K[1][1] = xc1;
    double d2 = 0;
    if(s == TRI) {
      for(int e = 0; e < 3; ++e) {
        int a = e, b = (e + 1) % 3;
        double t = 0;
        for(int k = 0; k < 3; ++k) t += (X[a][k] - X[b][k]) * (X[a][k] - X[b][k]);
        if(t > d2) d2 = t;
      }
    } else {
      for(int e = 0; e < 2; ++e) {
        double t = 0;
        for(int k = 0; k < 3; ++k) t += (X[e][k] - X[e + 2][k]) * (X[e][k] - X[e + 2][k]);
        if(t > d2) d2 = t;
      }
    }
    g->hK = sqrt(d2);
    return;
  }
  for(int a = 0; a < 3; ++a) {
    if(s == TET) {
      g->c0[a] = X[0][a];
      for(int k = 0; k < 3; ++k) g->C[a][k] = X[k + 1][a] - X[0][a];
    } else {
      g->c0[a] = (X[1][a] + X[3][a] + X[4][a] - X[0][a]) * 0.5;
      g->C[a][0] = (X[1][a] - X[0][a]) * 0.5;
      g->C[a][1] = (X[3][a] - X[0][a]) * 0.5;
      g->C[a][2] = (X[4][a] - X[0][a]) * 0.5;
    }
  }
  const double xc1 = g->C[0][0], xc2 = g->C[0][1], xc3 = g->C[0][2];
  const double yc1 = g->C[1][0], yc2 = g->C[1][1], yc3 = g->C[1][2];
  const double zc1 = g->C[2][0], zc2 = g->C[2][1], zc3 = g->C[2][2];
  g->det = xc1 * yc2 * zc3 + xc2 * yc3 * zc1 + xc3 * yc1 * zc2 - xc3 * yc2 * zc1 - xc2 * yc1 * zc3 -
           xc1 * yc3 * zc2;
  g->K[0][0] = yc2 * zc3 - yc3 * zc2; g->K[0][1] = yc3 * zc1 - yc1 * zc3; g->K[0][2] = yc1 * zc2 - yc2 * zc1;
  g->K[1][0] = xc3 * zc2 - xc2 * zc3; g->K[1][1] = xc1 * zc3 - xc3 * zc1; g->K[1][2] = xc2 * zc1 - xc1 * zc2;
  g->K[2][0] = xc2 * yc3 - xc3 * yc2; g->K[2][1] = xc3 * yc1 - xc1 * yc3; g->K[2][2] = xc1 * yc2 - xc2 * yc1;
  double d2 = 0;
  if(s == TET) {
    for(int a = 0; a < 4; ++a)
      for(int b = a + 1; b < 4; ++b) {
        double t = 0;
        for(int k = 0; k < 3; ++k) t += (X[a][k] - X[b][k]) * (X[a][k] - X[b][k]);
        if(t > d2) d2 = t;
      }
  } else {
    const int dg[4][2] = {{0, 6}, {1, 7}, {2, 4}, {3, 5}};
    for(int e = 0; e < 4; ++e) {
      double t = 0;
      for(int k = 0; k < 3; ++k) t += (X[dg[e][0]][k] - X[dg[e][1]][k]) * (X[dg[e][0]][k] - X[dg[e][1]][k]);
      if(t > d2) d2 = t;
    }
  }
  g->hK = sqrt(d2);
}

static void geo_point(const ogeo* g, int dim, const double* r, double* x) {
  for(int a = 0; a < 3; ++a) {
    x[a] = g->c0[a] + g->C[a][0] * r[0] + g->C[a][1] * r[1];
    if(dim == 3) x[a] = x[a] + g->C[a][2] * r[2];
  }
}

/* node position of local DOF i in cell c: affine image of the reference node */
static void node_xyz(const orc_problem* P, const omesh* m, int c, int i, double* x) {
  ogeo g;
  geo(m, c, &g);
  geo_point(&g, P->el.dim, P->el.node[i], x);
}

/* ----------------------------------------------------------------------- */
/* DOF numbering (union-find restatement)                                   */
/* ----------------------------------------------------------------------- */
#define OR_FIRSTMARK (-10)
#define NBT 7

static int uf_find(const int* G, int w) {
  int v;
  while((v = G[w]) > -1) w = v;
  return w;
}

static void map_single_dof(int* G, int dof0, int dof1, int* Counter) {
  int w0 = uf_find(G, dof0), w1 = uf_find(G, dof1);
  if(G[dof0] == -1 && G[dof1] == -1) {
    (*Counter)--;
    if(dof0 < dof1) { G[dof0] = *Counter; G[dof1] = dof0; }
    else { G[dof1] = *Counter; G[dof0] = dof1; }
  } else if(G[dof0] == -1) {
    if(w1 > dof0) { G[dof0] = G[w1]; G[dof1] = dof0; G[w1] = dof0; }
    else G[dof0] = w1;
  } else if(G[dof1] == -1) {
    if(w0 > dof1) { G[dof1] = G[w0]; G[dof0] = dof1; G[w0] = dof1; }
    else G[dof1] = w0;
  } else if(G[w0] != G[w1]) {
    int e = G[w0] > G[w1] ? G[w0] : G[w1];
    int w = w0 < w1 ? w0 : w1;
    G[w0] = w; G[w1] = w; G[w] = e;
  }
}

static void map_single_boundary_dof(int* G, int dof, int* BoundCounter) {
  int w = uf_find(G, dof);
  if(G[w] == -1) {
    (*BoundCounter)--;
    G[dof] = *BoundCounter;
  } else if(G[w] < *BoundCounter) {
    (*BoundCounter)--;
    G[w] = *BoundCounter;
    if(dof != w) G[dof] = w;
  }
}

/* local DOFs lying on joint j (nodes inside the joint) */
static int joint_dofs(const orc_problem* P, int j, int* out) {
  const oelem* e = &P->el;
  const int s = e->shape, len = jv_len(s);
  /* reference vertex coordinates */
  double rv[8][3];
  oelem e1 = make_elem(s, 1);
  for(int v = 0; v < NV[s]; ++v) {
    /* vertex v of the shape: for tensor shapes the element-1 node index of
       the corner is the tensor position of the vertex */
    if(s == QUAD || s == HEX) {
      static const int q2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
      rv[v][0] = q2[v % 4][0] ? 1 : -1;
      rv[v][1] = q2[v % 4][1] ? 1 : -1;
      rv[v][2] = s == HEX ? (v / 4 ? 1 : -1) : 0;
    } else {
      for(int k = 0; k < 3; ++k) rv[v][k] = e1.node[v][k];
    }
  }
  int n = 0;
  for(int i = 0; i < e->nl; ++i) {
    /* node inside the convex hull of the joint vertices: solve by checking
       it is an affine combination with non-negative weights (joint is a
       segment, triangle or parallelogram on the reference cell) */
    const double* x = e->node[i];
    int in = 0;
    if(len == 2) {
      const double* a = rv[jv(s, j, 0)];
      const double* b = rv[jv(s, j, 1)];
      double cr = (b[0] - a[0]) * (x[1] - a[1]) - (b[1] - a[1]) * (x[0] - a[0]);
      in = fabs(cr) < 1e-12;
    } else {
      const double* a = rv[jv(s, j, 0)];
      const double* b = rv[jv(s, j, 1)];
      const double* c = rv[jv(s, j, 2)];
      double u[3], v[3], w[3], nrm[3];
      for(int k = 0; k < 3; ++k) { u[k] = b[k] - a[k]; v[k] = c[k] - a[k]; w[k] = x[k] - a[k]; }
      nrm[0] = u[1] * v[2] - u[2] * v[1];
      nrm[1] = u[2] * v[0] - u[0] * v[2];
      nrm[2] = u[0] * v[1] - u[1] * v[0];
      in = fabs(nrm[0] * w[0] + nrm[1] * w[1] + nrm[2] * w[2]) < 1e-12;
    }
    if(in) out[n++] = i;
  }
  return n;
}

static void build_space(orc_problem* P, olevel* L) {
  const omesh* m = &L->m;
  const int s = m->shape, nj = NJ[s], nl = P->el.nl;
  const int N = m->ncells * nl;
  L->nl = nl;
  int jd[8][MAXNL], njd[8];
  for(int j = 0; j < nj; ++j) njd[j] = joint_dofs(P, j, jd[j]);
  /* BoundaryUpperBound: joint DOFs per boundary joint, by kind */
  int BUB[NBT] = {0};
  for(int c = 0; c < m->ncells; ++c)
    for(int j = 0; j < nj; ++j)
      if(L->nb_cell[c * nj + j] == -1) BUB[ex_bc(&P->ex, L->comp[c * nj + j])] += njd[j];
  int Counters[NBT + 1];
  Counters[0] = OR_FIRSTMARK + 1;
  for(int k = 0; k < NBT; ++k) Counters[k + 1] = Counters[k] - BUB[k];
  int* Counter = &Counters[NBT];
  int* G = malloc(sizeof(int) * N);
  for(int i = 0; i < N; ++i) G[i] = -1;
  /* nodal point coordinates of every (cell, local dof) */
  double* NX = malloc(sizeof(double) * 3 * (size_t)N);
  for(int c = 0; c < m->ncells; ++c) {
    ogeo g;
    geo(m, c, &g);
    for(int i = 0; i < nl; ++i) geo_point(&g, P->el.dim, P->el.node[i], &NX[3 * ((size_t)c * nl + i)]);
  }
  for(int c = 0; c < m->ncells; ++c) {
    for(int j = 0; j < nj; ++j) {
      const int nb = L->nb_cell[c * nj + j];
      if(nb < 0) {
        const int kind = ex_bc(&P->ex, L->comp[c * nj + j]);
        for(int k = 0; k < njd[j]; ++k) map_single_boundary_dof(G, c * nl + jd[j][k], &Counters[kind]);
      } else if(nb > c) {
        const int jn = L->nb_joint[c * nj + j];
        /* pair joint DOFs by coinciding nodal points */
        for(int k = 0; k < njd[j]; ++k) {
          const double* x0 = &NX[3 * ((size_t)c * nl + jd[j][k])];
          for(int kk = 0; kk < njd[jn]; ++kk) {
            const double* x1 = &NX[3 * ((size_t)nb * nl + jd[jn][kk])];
            if(fabs(x0[0] - x1[0]) + fabs(x0[1] - x1[1]) + fabs(x0[2] - x1[2]) < 1e-12) {
              map_single_dof(G, c * nl + jd[j][k], nb * nl + jd[jn][kk], Counter);
              break;
            }
          }
        }
      }
    }
    /* inner DOFs (on no joint) */
    for(int i = 0; i < nl; ++i) {
      int on = 0;
      for(int j = 0; j < nj && !on; ++j)
        for(int k = 0; k < njd[j]; ++k)
          if(jd[j][k] == i) on = 1;
      if(!on) { (*Counter)--; G[c * nl + i] = *Counter; }
    }
  }
  free(NX);
  /* compute_various_numbers (no hanging nodes) */
  int total = 0;
  for(int k = 0; k < NBT; ++k) total += BUB[k];
  int BoundMark[NBT];
  int l = OR_FIRSTMARK;
  for(int k = 0; k < NBT; ++k) { l -= BUB[k]; BoundMark[k] = l; }
  const int InnerMark = l;
  int BoundCounter[NBT];
  l = 0;
  for(int k = 0; k < NBT; ++k) { l += BUB[k]; BoundCounter[k] = l; }
  int count = l, DirichletCounter = 0, N_Dir = 0, N_Inner = 0;
  int NBnd[NBT] = {0};
  for(int i = 0; i < N; ++i) {
    int v = G[i];
    if(v < -1) {
      if(v <= InnerMark) { G[i] = count++; N_Inner++; continue; }
      int mm = NBT - 1, matched = 0;
      for(; mm >= 0; --mm)
        if(v <= BoundMark[mm]) { G[i] = BoundCounter[mm]++; NBnd[mm]++; matched = 1; break; }
      if(!matched && v <= OR_FIRSTMARK) { G[i] = DirichletCounter++; N_Dir++; }
    } else if(v >= 0) {
      G[i] = G[v];
    }
  }
  int nb_total = 0;
  for(int k = 0; k < NBT; ++k) nb_total += NBnd[k];
  const int DirichletOffset = N_Inner + nb_total;
  const int InnerOffset = -total;
  int BoundOffset[NBT];
  l = N_Inner;
  {
    int mm = 0;
    for(int k = 0; k < NBT; ++k) { mm += BUB[k]; BoundOffset[k] = l - mm; l += NBnd[k]; }
  }
  const int DirichletMark = BUB[0];
  l = BUB[0];
  for(int k = 0; k < NBT - 1; ++k) { l += BUB[k + 1]; BoundMark[k] = l; }
  BoundMark[NBT - 1] = l;
  for(int i = 0; i < N; ++i) {
    int v = G[i];
    if(v < DirichletMark) G[i] += DirichletOffset;
    else if(v < BoundMark[NBT - 1]) {
      for(int mm = 0; mm < NBT; ++mm)
        if(v < BoundMark[mm]) { G[i] += BoundOffset[mm]; break; }
    } else G[i] += InnerOffset;
  }
  L->dof = G;
  L->nactive = N_Inner + nb_total;
  L->ndir = N_Dir;
  L->ndof = L->nactive + N_Dir;
}

/* sorted full pattern (Structure.C) */
static int icmp(const void* a, const void* b) { return *(const int*)a - *(const int*)b; }
static void build_structure(olevel* L) {
  const int n = L->ndof, nl = L->nl, nc = L->m.ncells;
  int* cnt = calloc(n + 1, sizeof(int));
  for(int i = 0; i < nc * nl; ++i) cnt[L->dof[i] + 1]++;
  for(int r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
  int* cells = malloc(sizeof(int) * nc * nl);
  int* fill = malloc(sizeof(int) * n);
  memcpy(fill, cnt, sizeof(int) * n);
  for(int c = 0; c < nc; ++c)
    for(int m = 0; m < nl; ++m) cells[fill[L->dof[c * nl + m]]++] = c;
  L->rp = malloc(sizeof(int) * (n + 1));
  int* buf = malloc(sizeof(int) * 64 * MAXNL);
  long cap = (long)n * 32 + 64, used = 0;
  L->col = malloc(sizeof(int) * cap);
  L->rp[0] = 0;
  for(int r = 0; r < n; ++r) {
    int nb = 0;
    for(int e = cnt[r]; e < cnt[r + 1]; ++e)
      for(int k = 0; k < nl; ++k) buf[nb++] = L->dof[cells[e] * nl + k];
    qsort(buf, nb, sizeof(int), icmp);
    int u = 0;
    for(int k = 0; k < nb; ++k)
      if(u == 0 || buf[k] != buf[u - 1]) buf[u++] = buf[k];
    if(used + u > cap) { cap = (used + u) * 2; L->col = realloc(L->col, sizeof(int) * cap); }
    memcpy(&L->col[used], buf, sizeof(int) * u);
    used += u;
    L->rp[r + 1] = (int)used;
  }
  L->nnz = used;
  L->val = calloc(used, sizeof(double));
  L->rhs = calloc(n, sizeof(double));
  L->sol = calloc(n, sizeof(double));
  free(buf); free(cells); free(fill); free(cnt);
}

/* ---- spaces on meshes with hanging vertices (P1/Q1, P2/Q2) ----
   TFESpace::compute_various_numbers (src/FE/FESpace.C:1863-2099) orders the
   DOFs as inner, non-Dirichlet boundary kinds, hanging, Dirichlet; a DOF on
   boundary joints of several kinds takes the lowest kind (Dirichlet first,
   as map_single_boundary_dof keeps the later, larger mark). DOFs are
   identified through the mesh entity at their nodal point (vertex, edge,
   cell interior); within a class they are numbered in first appearance over
   (cell, local DOF), hanging DOFs in the order their coarse edges are met.
   A leaf edge (u, w) whose midpoint m exists is refined on the other side
   (FE2DMapper1Reg.C:43-78, the one-regular mappers):
   - P1/Q1 (C1_2_C1_2_1Reg): the vertex m hangs, coupled to u and w with
     HN_C_P1_2D_0 = {1/2, 1/2} (HNDesc.C:9-11);
   - P2/Q2 (C2_2_C2_2_1Reg): the coarse edge's middle DOF and the fine
     vertex m are one DOF; the middle DOFs of the fine edges (u, m) and
     (m, w), at 1/4 and 3/4 of the coarse edge, hang, coupled to the coarse
     edge's DOFs (u, m, w) with the quadratic Lagrange values there,
     {3/8, 3/4, -1/8} and {-1/8, 3/4, 3/8} (HN_C_P2_2D_1 and _0,
     HNDesc.C:12-17).
   sort_hanging_nodes (FESpace.C:2101-2160): depth-first post-order over the
   coupling DOFs that are hanging themselves, then reversed, so that a
   hanging DOF comes before the hanging DOFs it couples with. */
#define MAXCPL 3
static void hang_visit(const olevel* L, int h, char* seen, int* out, int* n) {
  seen[h] = 1;
  for(int k = 0; k < L->hang_ncpl[h]; ++k) {
    const int c = L->hang_cpl[MAXCPL * h + k] - L->nactive_nh;
    if(c >= 0 && c < L->nhang && !seen[c]) hang_visit(L, c, seen, out, n);
  }
  out[(*n)++] = h;
}

static void dict_init(vdict* d, int cap) {
  d->cap = cap;
  d->slot = malloc(sizeof(int) * cap);
  d->key = malloc(sizeof(int[4]) * cap);
  for(int i = 0; i < cap; ++i) d->slot[i] = -1;
}

/* value stored for the edge (u, w), or -1 */
static int edge_class(const vdict* d, int u, int w) {
  int key[4] = {u < w ? u : w, u < w ? w : u, -1, -1};
  return vertex_find(d, key);
}

static void edge_class_set(vdict* d, int u, int w, int v) {
  int key[4] = {u < w ? u : w, u < w ? w : u, -1, -1};
  uint64_t h = hash4(key) & (uint64_t)(d->cap - 1);
  while(d->slot[h] >= 0) h = (h + 1) & (uint64_t)(d->cap - 1);
  d->slot[h] = v;
  memcpy(d->key[h], key, sizeof(key));
}

static int build_space_hanging(orc_problem* P, olevel* L, const vdict* mids) {
  const omesh* m = &L->m;
  const int s = m->shape, nj = NJ[s], nv = NV[s], nl = P->el.nl, order = P->el.order;
  if(order != 1 && order != 2) return -1;
  L->nl = nl;
  /* entity of local DOF i: vertex a (etype 0), edge j (1) or interior (2) */
  int etype[MAXNL], eidx[MAXNL];
  {
    ogeo g;
    geo(m, 0, &g);
    for(int i = 0; i < nl; ++i) {
      double x[3];
      geo_point(&g, P->el.dim, P->el.node[i], x);
      etype[i] = 2;
      eidx[i] = -1;
      for(int a = 0; a < nv; ++a) {
        const double* X = &m->X[3 * m->cv[a]];
        if(fabs(X[0] - x[0]) + fabs(X[1] - x[1]) + fabs(X[2] - x[2]) < 1e-12) { etype[i] = 0; eidx[i] = a; }
      }
      for(int j = 0; j < nj && etype[i] == 2; ++j) {
        const double* A = &m->X[3 * m->cv[jv(s, j, 0)]];
        const double* B = &m->X[3 * m->cv[jv(s, j, 1)]];
        double d = 0;
        for(int k = 0; k < 3; ++k) d += fabs(0.5 * (A[k] + B[k]) - x[k]);
        if(d < 1e-12) { etype[i] = 1; eidx[i] = j; }
      }
    }
  }
  /* classes in first appearance */
  const int NVt = m->nverts, N = m->ncells * nl;
  int* vcls = malloc(sizeof(int) * NVt);
  for(int v = 0; v < NVt; ++v) vcls[v] = -1;
  vdict ed;
  {
    int cap = 1;
    while(cap < 4 * (m->ncells * nj + 16)) cap <<= 1;
    dict_init(&ed, cap);
  }
  int* cls = malloc(sizeof(int) * N);
  int ncls = 0;
  for(int c = 0; c < m->ncells; ++c)
    for(int i = 0; i < nl; ++i) {
      int k;
      if(etype[i] == 0) {
        const int v = m->cv[c * nv + eidx[i]];
        if(vcls[v] < 0) vcls[v] = ncls++;
        k = vcls[v];
      } else if(etype[i] == 1) {
        const int u = m->cv[c * nv + jv(s, eidx[i], 0)], w = m->cv[c * nv + jv(s, eidx[i], 1)];
        const int mm = edge_mid(mids, u, w);
        if(mm >= 0) { /* the coarse edge's middle DOF is the fine vertex */
          if(vcls[mm] < 0) vcls[mm] = ncls++;
          k = vcls[mm];
        } else {
          k = edge_class(&ed, u, w);
          if(k < 0) { k = ncls++; edge_class_set(&ed, u, w, k); }
        }
      } else {
        k = ncls++;
      }
      cls[c * nl + i] = k;
    }
  int* kind = malloc(sizeof(int) * ncls); /* NBT + 1 inner, NBT hanging, else boundary kind */
  for(int k = 0; k < ncls; ++k) kind[k] = NBT + 1;
  for(int c = 0; c < m->ncells; ++c)
    for(int j = 0; j < nj; ++j) {
      if(L->nb_cell[c * nj + j] != -1) continue;
      const int b = ex_bc(&P->ex, L->comp[c * nj + j]);
      int jd[MAXNL];
      const int n = joint_dofs(P, j, jd);
      for(int q = 0; q < n; ++q) {
        const int k = cls[c * nl + jd[q]];
        if(b < kind[k]) kind[k] = b;
      }
    }
  /* hanging classes, couplings and weights in the order the coarse edges are met */
  int cap_h = 2 * (m->ncells * nj + 1);
  int* hcls = malloc(sizeof(int) * cap_h);
  int* hcpl = malloc(sizeof(int) * MAXCPL * cap_h);
  double* hw = malloc(sizeof(double) * MAXCPL * cap_h);
  int* hn = malloc(sizeof(int) * cap_h);
  char* met = calloc(NVt, 1);
  int nh = 0;
  for(int c = 0; c < m->ncells; ++c)
    for(int j = 0; j < nj; ++j) {
      if(L->nb_cell[c * nj + j] != -2) continue;
      const int u = m->cv[c * nv + jv(s, j, 0)], w = m->cv[c * nv + jv(s, j, 1)];
      const int mm = edge_mid(mids, u, w);
      if(mm < 0 || met[mm]) continue;
      met[mm] = 1;
      if(order == 1) {
        hcls[nh] = vcls[mm];
        hn[nh] = 2;
        hcpl[MAXCPL * nh] = vcls[u]; hcpl[MAXCPL * nh + 1] = vcls[w];
        hw[MAXCPL * nh] = 0.5; hw[MAXCPL * nh + 1] = 0.5;
        kind[hcls[nh]] = NBT;
        nh++;
      } else {
        const int half[2][2] = {{u, mm}, {mm, w}};
        const double wt[2][3] = {{0.375, 0.75, -0.125}, {-0.125, 0.75, 0.375}};
        for(int q = 0; q < 2; ++q) {
          const int k = edge_class(&ed, half[q][0], half[q][1]);
          if(k < 0) { free(cls); free(kind); free(vcls); free(hcls); free(hcpl); free(hw); free(hn); free(met);
                      free(ed.slot); free(ed.key); return -1; }
          hcls[nh] = k;
          hn[nh] = 3;
          hcpl[MAXCPL * nh] = vcls[u]; hcpl[MAXCPL * nh + 1] = vcls[mm]; hcpl[MAXCPL * nh + 2] = vcls[w];
          for(int t = 0; t < 3; ++t) hw[MAXCPL * nh + t] = wt[q][t];
          kind[k] = NBT;
          nh++;
        }
      }
    }
  int* num = malloc(sizeof(int) * ncls);
  for(int k = 0; k < ncls; ++k) num[k] = -1;
  int next = 0;
  /* inner, then non-Dirichlet boundary kinds 1..NBT-1 */
  for(int o = 0; o < NBT; ++o) {
    const int want = o == 0 ? NBT + 1 : o;
    for(int p = 0; p < N; ++p)
      if(kind[cls[p]] == want && num[cls[p]] < 0) num[cls[p]] = next++;
  }
  L->nactive_nh = next;
  for(int h = 0; h < nh; ++h) num[hcls[h]] = next++;
  L->nactive = next;
  for(int p = 0; p < N; ++p)
    if(kind[cls[p]] == 0 && num[cls[p]] < 0) num[cls[p]] = next++;
  L->ndof = next;
  L->ndir = L->ndof - L->nactive;
  L->dof = malloc(sizeof(int) * N);
  for(int p = 0; p < N; ++p) L->dof[p] = num[cls[p]];
  L->nhang = nh;
  L->hang_ncpl = malloc(sizeof(int) * (nh + 1));
  L->hang_cpl = malloc(sizeof(int) * MAXCPL * (nh + 1));
  L->hang_w = malloc(sizeof(double) * MAXCPL * (nh + 1));
  for(int h = 0; h < nh; ++h) {
    L->hang_ncpl[h] = hn[h];
    for(int t = 0; t < hn[h]; ++t) {
      L->hang_cpl[MAXCPL * h + t] = num[hcpl[MAXCPL * h + t]];
      L->hang_w[MAXCPL * h + t] = hw[MAXCPL * h + t];
    }
  }
  L->hang_order = malloc(sizeof(int) * (nh + 1));
  char* seen = calloc(nh + 1, 1);
  int n = 0;
  for(int h = 0; h < nh; ++h)
    if(!seen[h]) hang_visit(L, h, seen, L->hang_order, &n);
  for(int a = 0, b = n - 1; a < b; ++a, --b) { int t = L->hang_order[a]; L->hang_order[a] = L->hang_order[b]; L->hang_order[b] = t; }
  free(seen); free(num); free(cls); free(kind); free(vcls); free(hcls); free(hcpl); free(hw); free(hn); free(met);
  free(ed.slot); free(ed.key);
  return 0;
}

/* rows of the coupling DOFs take the columns of the hanging rows added to
   them (ModifyMatrixAccordingToCoupling needs every entry it adds,
   src/Matrix/FEMatrix.C:244-296); the pattern of Structure.C:75-230 gives
   them to every coupling row, Dirichlet rows included, where they stay
   zero. (For P1 triangles the reference also adds the transposed couplings,
   Structure.C:196-218, zeros that only its AFC uses; they are left out.) */
static void hanging_structure(olevel* L) {
  const int n = L->ndof;
  int** rows = malloc(sizeof(int*) * n);
  int* len = malloc(sizeof(int) * n);
  for(int r = 0; r < n; ++r) {
    len[r] = L->rp[r + 1] - L->rp[r];
    rows[r] = malloc(sizeof(int) * (len[r] + MAXCPL));
    memcpy(rows[r], &L->col[L->rp[r]], sizeof(int) * len[r]);
  }
  /* a hanging row holds its coupling DOFs, also those outside its cells
     (add_columns_far_hanging_coupling, Structure.C:142-170: P2 and up) */
  for(int h = 0; h < L->nhang; ++h) {
    const int hd = L->nactive_nh + h;
    for(int k = 0; k < L->hang_ncpl[h]; ++k) {
      const int c = L->hang_cpl[MAXCPL * h + k];
      int pos = 0;
      while(pos < len[hd] && rows[hd][pos] < c) pos++;
      if(pos < len[hd] && rows[hd][pos] == c) continue;
      memmove(&rows[hd][pos + 1], &rows[hd][pos], sizeof(int) * (len[hd] - pos));
      rows[hd][pos] = c;
      len[hd]++;
    }
  }
  for(int t = 0; t < L->nhang; ++t) {
    const int h = L->hang_order[t], hd = L->nactive_nh + h;
    for(int k = 0; k < L->hang_ncpl[h]; ++k) {
      const int c = L->hang_cpl[MAXCPL * h + k];
      int* merged = malloc(sizeof(int) * (len[c] + len[hd]));
      int a = 0, b = 0, u = 0;
      while(a < len[c] || b < len[hd]) {
        int v;
        if(b >= len[hd] || (a < len[c] && rows[c][a] < rows[hd][b])) v = rows[c][a++];
        else if(a >= len[c] || rows[hd][b] < rows[c][a]) v = rows[hd][b++];
        else { v = rows[c][a++]; b++; }
        merged[u++] = v;
      }
      free(rows[c]);
      rows[c] = merged;
      len[c] = u;
    }
  }
  long total = 0;
  for(int r = 0; r < n; ++r) total += len[r];
  free(L->col);
  L->col = malloc(sizeof(int) * total);
  L->rp[0] = 0;
  for(int r = 0; r < n; ++r) {
    memcpy(&L->col[L->rp[r]], rows[r], sizeof(int) * len[r]);
    L->rp[r + 1] = L->rp[r] + len[r];
    free(rows[r]);
  }
  free(rows); free(len);
  L->nnz = total;
  free(L->val);
  L->val = calloc(total, sizeof(double));
}

/* Assemble2D.C:3350-3427 after the cell loop: for every hanging DOF in the
   sorted order, its row is added with weight 1/2 to the rows of its active
   (hanging included) coupling DOFs (ModifyMatrixAccordingToCoupling, with
   bound n_active), so is its rhs entry, then the hanging rhs entries are 0
   and the hanging rows become u_h - u_a/2 - u_b/2 = 0 (correct_hanging_rows,
   FEMatrix.C:384-418) */
static double* entry_at(const olevel* L, int r, int col) {
  for(int e = L->rp[r]; e < L->rp[r + 1]; ++e)
    if(L->col[e] == col) return &L->val[e];
  return NULL;
}

static void apply_hanging(olevel* L) {
  /* planted mistakes: 6 = hanging coupling DOFs receive nothing (bound
     n_active_non_hanging, as Assembler4 uses), 7 = list order, not sorted */
  const int bound = orc_fault == 6 ? L->nactive_nh : L->nactive;
  for(int t = 0; t < L->nhang; ++t) {
    const int h = orc_fault == 7 ? t : L->hang_order[t], hd = L->nactive_nh + h;
    for(int k = 0; k < L->hang_ncpl[h]; ++k) {
      const int c = L->hang_cpl[MAXCPL * h + k];
      const double w = L->hang_w[MAXCPL * h + k];
      if(c >= bound) continue;
      for(int e = L->rp[hd]; e < L->rp[hd + 1]; ++e) {
        double* dst = entry_at(L, c, L->col[e]);
        if(!dst) { fprintf(stderr, "oracle: hanging entry missing\n"); continue; }
        *dst += w * L->val[e];
      }
    }
  }
  for(int t = 0; t < L->nhang; ++t) {
    const int h = orc_fault == 7 ? t : L->hang_order[t], hd = L->nactive_nh + h;
    for(int k = 0; k < L->hang_ncpl[h]; ++k) {
      const int c = L->hang_cpl[MAXCPL * h + k];
      if(c < bound) L->rhs[c] += L->hang_w[MAXCPL * h + k] * L->rhs[hd];
    }
  }
  for(int h = 0; h < L->nhang; ++h) L->rhs[L->nactive_nh + h] = 0;
  for(int h = 0; h < L->nhang; ++h) {
    const int hd = L->nactive_nh + h;
    for(int e = L->rp[hd]; e < L->rp[hd + 1]; ++e) {
      const int col = L->col[e];
      double v = col == hd ? 1.0 : 0.0;
      for(int k = 0; k < L->hang_ncpl[h]; ++k)
        if(col == L->hang_cpl[MAXCPL * h + k]) v = -L->hang_w[MAXCPL * h + k];
      L->val[e] = v;
    }
  }
}

/* ----------------------------------------------------------------------- */
/* assembly                                                                 */
/* ----------------------------------------------------------------------- */
static double sdfem_delta(int dim, double hK, double eps, double b1, double b2, double b3) {
  if(fabs(eps) < 1e-20) eps = 1e-20;
  double nb = b1 * b1 + b2 * b2;
  if(dim == 3) nb += b3 * b3;
  nb = sqrt(nb);
  double delta;
  const double nu = 1.0 / 1;
  if(nb > 0) {
    double alpha = nu * nb * hK / (2 * eps);
    delta = nu * hK * (1 / tanh(alpha) - 1 / alpha) / (2 * nb);
  } else delta = 0;
  return delta * 1.0;
}


/* physical values at quad point: u, ux, uy, uz, uxx, uyy, uzz for all i */
static void orig_values(const orc_problem* P, const ogeo* g, const double* ref, double (*o)[MAXNL]) {
  const int nl = P->el.nl, d = P->el.dim, s = P->el.shape;
  for(int i = 0; i < nl; ++i) {
    o[0][i] = ref[i];
    if(d == 2) {
      const double rx = ref[nl + i], ry = ref[2 * nl + i];
      o[1][i] = (g->K[0][0] * rx + g->K[0][1] * ry) / g->det;
      o[2][i] = (g->K[1][0] * rx + g->K[1][1] * ry) / g->det;
      o[3][i] = 0;
      for(int a = 0; a < 2; ++a) {
        const double j0 = g->K[a][0] / g->det, j1 = g->K[a][1] / g->det;
        o[4 + a][i] = j0 * j0 * ref[3 * nl + i] + (orc_fault == 2 ? 0.0 : 2 * j0 * j1 * ref[4 * nl + i]) +
                      j1 * j1 * ref[5 * nl + i];
      }
      o[6][i] = 0;
    } else {
      const double rx = ref[nl + i], ry = ref[2 * nl + i], rz = ref[3 * nl + i];
      for(int a = 0; a < 3; ++a) {
        if(s == TET)
          o[1 + a][i] = (g->K[a][0] / g->det) * rx + (g->K[a][1] / g->det) * ry + (g->K[a][2] / g->det) * rz;
        else
          o[1 + a][i] = (g->K[a][0] * rx + g->K[a][1] * ry + g->K[a][2] * rz) / g->det;
      }
      for(int a = 0; a < 3; ++a) {
        const double j0 = g->K[a][0] / g->det, j1 = g->K[a][1] / g->det, j2 = g->K[a][2] / g->det;
        const double mixed = orc_fault == 2 ? 0.0 : 1.0;
        o[4 + a][i] = j0 * j0 * ref[4 * nl + i] + mixed * 2 * j0 * j1 * ref[5 * nl + i] +
                      mixed * 2 * j0 * j2 * ref[6 * nl + i] + j1 * j1 * ref[7 * nl + i] +
                      mixed * 2 * j1 * j2 * ref[8 * nl + i] + j2 * j2 * ref[9 * nl + i];
      }
    }
  }
}

static void assemble_level(orc_problem* P, olevel* L, int c_begin, int c_end) {
  const omesh* m = &L->m;
  const int nl = L->nl, d = P->el.dim, s = m->shape, nj = NJ[s];
  const oquad* q = &P->qd;
  /* time-dependent forms: stiffness into aval, mass into mval */
  double* AV = P->tcd ? L->aval : L->val;
  memset(AV, 0, sizeof(double) * L->nnz);
  if(P->tcd) memset(L->mval, 0, sizeof(double) * L->nnz);
  memset(L->rhs, 0, sizeof(double) * L->ndof);
  static double ref[MAXQ][10 * MAXNL];
  for(int k = 0; k < q->n; ++k) elem_eval(&P->el, q->x[k], ref[k]);
  double A[MAXNL][MAXNL], F[MAXNL], Mm[MAXNL][MAXNL];
  double o[7][MAXNL];
  for(int c = c_begin; c < c_end; ++c) {
    ogeo g;
    geo(m, c, &g);
    memset(A, 0, sizeof(A));
    memset(F, 0, sizeof(F));
    memset(Mm, 0, sizeof(Mm));
    for(int k = 0; k < q->n; ++k) {
      double x[3], cf[6];
      geo_point(&g, d, q->x[k], x);
      const double Mult = q->w[k] * fabs(g.det);
      ex_coeffs(&P->ex, x[0], x[1], d == 3 ? x[2] : 0.0, cf);
      orig_values(P, &g, ref[k], o);
      const double eps = cf[0], b1 = cf[1], b2 = cf[2], b3 = cf[3], cr = cf[4], f = cf[5];
      const double delta = P->supg ? sdfem_delta(d, g.hK, eps, b1, b2, b3) : 0;
      if(P->tcd) {
        /* TCDStiff, TCDMass, TCDRhs, then TCDStiffSUPG, TCDMassSUPG,
           TCDRhsSUPG (CD_local_assembling_routines.C:9-387; SDFEM type 2:
           the parameter is not scaled with the time step) */
        for(int i = 0; i < nl; ++i) {
          const double t0 = o[0][i], tx = o[1][i], ty = o[2][i], tz = d == 3 ? o[3][i] : 0;
          F[i] += Mult * t0 * f;
          for(int j = 0; j < nl; ++j) {
            const double a0 = o[0][j], ax = o[1][j], ay = o[2][j], az = d == 3 ? o[3][j] : 0;
            double val = eps * (tx * ax + ty * ay + tz * az);
            val += (b1 * ax + b2 * ay + b3 * az) * t0;
            val += cr * t0 * a0;
            A[i][j] += Mult * val;
            Mm[i][j] += Mult * t0 * a0;
          }
          if(P->supg) {
            const double bgradv = delta * (b1 * tx + b2 * ty + b3 * tz);
            for(int j = 0; j < nl; ++j) {
              const double a0 = o[0][j], ax = o[1][j], ay = o[2][j], az = d == 3 ? o[3][j] : 0;
              const double lap = d == 3 ? o[4][j] + o[5][j] + o[6][j] : o[4][j] + o[5][j];
              double v2 = (b1 * ax + b2 * ay + b3 * az + cr * a0) * bgradv;
              v2 -= eps * lap * bgradv;
              A[i][j] += Mult * v2;
              Mm[i][j] += Mult * a0 * bgradv;
            }
            F[i] += Mult * f * bgradv;
          }
        }
        continue;
      }
      for(int i = 0; i < nl; ++i) {
        const double t0 = o[0][i], tx = o[1][i], ty = o[2][i], tz = o[3][i];
        /* Galerkin, then SD (the reference calls the local routines in this order) */
        if(d == 3) F[i] += Mult * t0 * f;
        else F[i] += Mult * t0 * f;
        for(int j = 0; j < nl; ++j) {
          const double a0 = o[0][j], ax = o[1][j], ay = o[2][j], az = o[3][j];
          if(d == 3) {
            double val = eps * (tx * ax + ty * ay + tz * az);
            val += (b1 * ax + b2 * ay + b3 * az) * t0;
            val += cr * a0 * t0;
            A[i][j] += val * Mult;
          } else {
            double val = eps * (tx * ax + ty * ay);
            val += (b1 * ax + b2 * ay) * t0;
            val += cr * a0 * t0;
            val *= Mult;
            A[i][j] += val;
          }
        }
      }
      if(P->supg) {
        for(int i = 0; i < nl; ++i) {
          const double tx = o[1][i], ty = o[2][i], tz = o[3][i];
          if(d == 3) {
            const double bgradv = b1 * tx + b2 * ty + b3 * tz;
            F[i] += Mult * (delta * bgradv) * f;
            for(int j = 0; j < nl; ++j) {
              double val = orc_fault == 1 ? 0.0 : -eps * (o[4][j] + o[5][j] + o[6][j]);
              val += b1 * o[1][j] + b2 * o[2][j] + b3 * o[3][j] + cr * o[0][j];
              A[i][j] += val * delta * bgradv * Mult;
            }
          } else {
            const double bgradv = b1 * tx + b2 * ty;
            F[i] += Mult * delta * bgradv * f;
            for(int j = 0; j < nl; ++j) {
              const double lap = orc_fault == 1 ? 0.0 : o[4][j] + o[5][j];
              double val = delta * (-eps * lap + b1 * o[1][j] + b2 * o[2][j] + cr * o[0][j]) * bgradv;
              val *= Mult;
              A[i][j] += val;
            }
          }
        }
      }
    }
    /* scatter by linear row search; Dirichlet rows get set(l,l,1) */
    const int* DOF = &L->dof[c * nl];
    for(int mm = 0; mm < nl; ++mm) {
      const int l = DOF[mm];
      if(l < L->nactive || P->afc) {
        for(int k = 0; k < nl; ++k) {
          int found = 0;
          for(int e = L->rp[l]; e < L->rp[l + 1]; ++e)
            if(L->col[e] == DOF[k]) {
              AV[e] += A[mm][k];
              if(P->tcd) L->mval[e] += Mm[mm][k];
              found = 1;
              break;
            }
          if(!found) fprintf(stderr, "oracle: entry not found\n");
        }
      } else {
        for(int e = L->rp[l]; e < L->rp[l + 1]; ++e)
          if(L->col[e] == l) {
            AV[e] = 1.0;
            if(P->tcd) L->mval[e] = 1.0;
          }
      }
    }
    for(int mm = 0; mm < nl; ++mm)
      if(DOF[mm] < L->nactive) L->rhs[DOF[mm]] += F[mm];
    /* boundary joints: Dirichlet values, Neumann integrals (2D) */
    for(int j = 0; j < nj; ++j) {
      if(L->nb_cell[c * nj + j] != -1) continue;
      const int comp = L->comp[c * nj + j];
      int jd[MAXNL];
      const int n = joint_dofs(P, j, jd);
      if(ex_bc(&P->ex, comp) == 0) {
        for(int k = 0; k < n; ++k) {
          double x[3];
          node_xyz(P, m, c, jd[k], x);
          L->rhs[DOF[jd[k]]] = ex_bvalue(&P->ex, comp, x[0], x[1], x[2]);
        }
      } else if(d == 2) {
        /* Gauss rule of degree 2p on the edge from vertex j to vertex j+1 */
        const int deg = 2 * P->el.order;
        double lw[4], lx[4];
        int ln;
        if(deg <= 3) { ln = 2; lw[0] = lw[1] = 1; lx[0] = 2 * 0.21132486540518711774 - 1; lx[1] = 2 * 0.78867513459481288226 - 1; }
        else {
          ln = 3;
          lw[0] = 0.2777777777777778 * 2; lw[1] = 0.4444444444444444 * 2; lw[2] = 0.2777777777777778 * 2;
          lx[0] = 2 * 0.11270166537925831149 - 1; lx[1] = 2 * 0.5 - 1; lx[2] = 2 * 0.88729833462074168851 - 1;
        }
        const int v0 = m->cv[c * NV[s] + jv(s, j, 0)], v1 = m->cv[c * NV[s] + jv(s, j, 1)];
        const double* x0 = &m->X[3 * v0];
        const double* x1 = &m->X[3 * v1];
        const double hE = sqrt((x1[0] - x0[0]) * (x1[0] - x0[0]) + (x1[1] - x0[1]) * (x1[1] - x0[1])) / 2;
        /* reference end points of the edge */
        double r0[3], r1[3];
        {
          oelem e1 = make_elem(s, 1);
          int a0 = jv(s, j, 0), a1 = jv(s, j, 1);
          if(s == QUAD) {
            static const int q2[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            r0[0] = q2[a0][0] ? 1 : -1; r0[1] = q2[a0][1] ? 1 : -1; r0[2] = 0;
            r1[0] = q2[a1][0] ? 1 : -1; r1[1] = q2[a1][1] ? 1 : -1; r1[2] = 0;
          } else {
            memcpy(r0, e1.node[a0], 24);
            memcpy(r1, e1.node[a1], 24);
          }
        }
        for(int lq = 0; lq < ln; ++lq) {
          const double t = (1 + lx[lq]) / 2;
          double rp[3], xp[3], phi[10 * MAXNL];
          for(int k = 0; k < 3; ++k) { rp[k] = r0[k] + t * (r1[k] - r0[k]); xp[k] = x0[k] + t * (x1[k] - x0[k]); }
          elem_eval(&P->el, rp, phi);
          double sv = ex_bvalue(&P->ex, comp, xp[0], xp[1], 0);
          sv *= hE * lw[lq];
          for(int k = 0; k < nl; ++k)
            if(DOF[k] < L->nactive) L->rhs[DOF[k]] += sv * phi[k];
        }
      }
    }
  }
}

/* ----------------------------------------------------------------------- */
/* linear algebra                                                           */
/* ----------------------------------------------------------------------- */
static double ddot(int n, const double* x, const double* y) {
  double r = 0;
  for(int i = 0; i < n; ++i) r += x[i] * y[i];
  return r;
}

/* y += a * A x on active rows, y[na] += a * x[na] */
static void apply_scaled_add(const olevel* L, const double* x, double* y, double a) {
  for(int i = 0; i < L->nactive; ++i) {
    double v = 0;
    for(int j = L->rp[i]; j < L->rp[i + 1]; ++j) v += L->val[j] * x[L->col[j]];
    y[i] += a * v;
  }
  for(int i = L->nactive; i < L->ndof; ++i) y[i] += a * x[i];
}

static void apply(const olevel* L, const double* x, double* y) {
  memset(y, 0, sizeof(double) * L->ndof);
  apply_scaled_add(L, x, y, 1.0);
}

static double diag_of(const olevel* L, int r) {
  for(int j = L->rp[r]; j < L->rp[r + 1]; ++j)
    if(L->col[j] == r) return L->val[j];
  return 0;
}

/* dense LU with partial pivoting of the level matrix (all rows as stored) */
static void dense_factor(olevel* L) {
  const int n = L->ndof;
  free(L->lu);
  free(L->piv);
  L->lu = calloc((size_t)n * n, sizeof(double));
  L->piv = malloc(sizeof(int) * n);
  for(int r = 0; r < n; ++r)
    for(int j = L->rp[r]; j < L->rp[r + 1]; ++j) L->lu[(size_t)r * n + L->col[j]] = L->val[j];
  double* a = L->lu;
  for(int k = 0; k < n; ++k) {
    int p = k;
    for(int i = k + 1; i < n; ++i)
      if(fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)p * n + k])) p = i;
    L->piv[k] = p;
    if(p != k)
      for(int j = 0; j < n; ++j) { double t = a[(size_t)k * n + j]; a[(size_t)k * n + j] = a[(size_t)p * n + j]; a[(size_t)p * n + j] = t; }
    for(int i = k + 1; i < n; ++i) {
      double f = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = f;
      if(f != 0)
        for(int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= f * a[(size_t)k * n + j];
    }
  }
}

static void dense_solve(const olevel* L, const double* b, double* x) {
  const int n = L->ndof;
  const double* a = L->lu;
  memcpy(x, b, sizeof(double) * n);
  for(int k = 0; k < n; ++k)
    if(L->piv[k] != k) { double t = x[k]; x[k] = x[L->piv[k]]; x[L->piv[k]] = t; }
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < i; ++j) x[i] -= a[(size_t)i * n + j] * x[j];
  for(int i = n - 1; i >= 0; --i) {
    for(int j = i + 1; j < n; ++j) x[i] -= a[(size_t)i * n + j] * x[j];
    x[i] /= a[(size_t)i * n + i];
  }
}

/* ----------------------------------------------------------------------- */
/* grid transfer (GridTransfer.C restatement, cell-wise with averaging)     */
/* ----------------------------------------------------------------------- */
/* coarse basis values at the fine nodes of child cell fc (QQ of the reference) */
static void child_matrix(const orc_problem* P, const olevel* C, const olevel* F, int fc, double* QQ) {
  const int pc = F->m.parent[fc];
  ogeo g;
  geo(&C->m, pc, &g);
  const int nl = P->el.nl, d = P->el.dim;
  double phi[10 * MAXNL];
  for(int k = 0; k < nl; ++k) {
    double x[3];
    node_xyz(P, &F->m, fc, k, x);
    /* reference coordinates in the parent: solve C r = x - c0 */
    double rhs[3] = {x[0] - g.c0[0], x[1] - g.c0[1], x[2] - g.c0[2]}, r[3] = {0, 0, 0};
    for(int a = 0; a < d; ++a) {
      double s = 0;
      for(int b = 0; b < d; ++b) s += g.K[b][a] * rhs[b];
      r[a] = s / g.det;
    }
    elem_eval(&P->el, r, phi);
    for(int l = 0; l < nl; ++l) QQ[k * nl + l] = phi[l];
  }
}

/* QQ depends only on the element and the child number on uniformly refined
   affine meshes (the reference caches it per (element, refinement, child),
   FEDatabase::GetProlongationMatrix3D); children of a cell are stored
   consecutively, so the child number is fc % n_children. */
static const double* child_qq(const orc_problem* P, const olevel* C, olevel* F, int fc) {
  const int nch = F->m.ncells / C->m.ncells, nl = P->el.nl;
  const int j = fc % nch;
  if(!F->qq) {
    F->qq = malloc(sizeof(double) * 8 * nl * nl);
    for(int k = 0; k < 8; ++k) F->qq_ready[k] = 0;
  }
  if(!F->qq_ready[j]) {
    child_matrix(P, C, F, fc, &F->qq[j * nl * nl]);
    F->qq_ready[j] = 1;
  }
  return &F->qq[j * nl * nl];
}

static void prolongate(const orc_problem* P, const olevel* C, olevel* F, const double* cf, double* ff) {
  const int nl = P->el.nl;
  double* aux = calloc(F->ndof, sizeof(double));
  memset(ff, 0, sizeof(double) * F->ndof);
  for(int i = 0; i < F->m.ncells; ++i)
    for(int j = 0; j < nl; ++j) aux[F->dof[i * nl + j]] += 1;
  double Val[MAXNL];
  for(int fc = 0; fc < F->m.ncells; ++fc) {
    const int pc = F->m.parent[fc];
    const double* QQ = child_qq(P, C, F, fc);
    for(int l = 0; l < nl; ++l) Val[l] = cf[C->dof[pc * nl + l]];
    for(int k = 0; k < nl; ++k) {
      double s = 0;
      for(int l = 0; l < nl; ++l) s += QQ[k * nl + l] * Val[l];
      ff[F->dof[fc * nl + k]] += s;
    }
  }
  for(int i = 0; i < F->ndof; ++i) ff[i] /= aux[i];
  free(aux);
}

static void restrict_defect(const orc_problem* P, const olevel* C, olevel* F, double* cf, const double* ff) {
  const int nl = P->el.nl;
  double* aux = calloc(F->ndof, sizeof(double));
  double* copy = malloc(sizeof(double) * F->ndof);
  memcpy(copy, ff, sizeof(double) * F->ndof);
  memset(cf, 0, sizeof(double) * C->ndof);
  for(int i = 0; i < F->m.ncells; ++i)
    for(int j = 0; j < nl; ++j) aux[F->dof[i * nl + j]] += 1;
  for(int i = 0; i < F->ndof; ++i) copy[i] /= aux[i];
  double Val[MAXNL];
  for(int fc = 0; fc < F->m.ncells; ++fc) {
    const int pc = F->m.parent[fc];
    const double* QQ = child_qq(P, C, F, fc);
    for(int l = 0; l < nl; ++l) Val[l] = copy[F->dof[fc * nl + l]];
    for(int k = 0; k < nl; ++k) {
      double s = 0;
      for(int l = 0; l < nl; ++l) s += QQ[l * nl + k] * Val[l];
      cf[C->dof[pc * nl + k]] += s;
    }
  }
  free(aux);
  free(copy);
}

/* ----------------------------------------------------------------------- */
/* solvers                                                                  */
/* ----------------------------------------------------------------------- */
typedef struct {
  int method, precond, max_iter, min_iter;
  double res_tol, res_red, div_factor;
  int restart, n_pre, n_post;
  double jacobi_damp, mg_damp;
  int mixed;
  int smoother; /* multigrid_smoother: 0 jacobi, 1 sor, 2 ssor */
  double sor_omega;
  int cycle; /* multigrid_cycle_type: 0 V, 1 W, 2 F */
  int coarse_smoother; /* multigrid_smoother_coarse: 0 direct solve, 1 jacobi, 2 sor, 3 ssor */
  double coarse_residual;
  int coarse_max_iter;
} osolver;

typedef struct {
  int iterations, converged;
  double initial_residual, final_residual, seconds;
} ostats;

typedef struct {
  const osolver* s;
  double initial, previous;
  int diverged;
} oconv;

static double g_hist[100000];
static int g_nhist;
static int converged(oconv* c, double res, int it) {
  if(it == 0) { c->initial = res; c->previous = res; g_nhist = 0; }
  if(g_nhist < 100000) g_hist[g_nhist++] = res;
  int small = res < c->s->res_tol, red = res < c->s->res_red * c->initial;
  int enough = it >= c->s->min_iter;
  int stag = res > 1e10 * c->previous, div = res > c->s->div_factor * c->initial;
  int conv = (small || red || stag || div) && enough;
  if(it == c->s->max_iter) conv = 1;
  if(div) c->diverged = 1;
  c->previous = res;
  return conv;
}

static void jacobi_smooth(const olevel* L, const double* b, double* x, double omega) {
  const int n = L->ndof;
  double* d = malloc(sizeof(double) * n);
  memcpy(d, b, sizeof(double) * n);
  apply_scaled_add(L, x, d, -1.0);
  for(int i = 0; i < n; ++i) d[i] /= L->diag[i];
  for(int i = 0; i < n; ++i) x[i] += omega * d[i];
  free(d);
}

/* TMatrix::sor_sweep (Matrix.C:1308-1420) on a CSR matrix, in place, over
   all rows in serial order: forward (flag 0), backward (1) or both (2). Each
   row sums b - A x in column order, then x += omega * s / diag. */
static void sor_csr(int n, const int* rp, const int* col, const double* val, const double* b, double* x,
                    double omega, int flag) {
  double* d = malloc(sizeof(double) * n);
  for(int r = 0; r < n; ++r) {
    d[r] = 0;
    for(int k = rp[r]; k < rp[r + 1]; ++k)
      if(col[k] == r) d[r] = val[k];
  }
  for(int pass = 0; pass < 2; ++pass) {
    const int backward = pass == 1;
    if(backward && flag == 0) break;
    if(!backward && flag == 1) continue;
    for(int t = 0; t < n; ++t) {
      const int r = backward ? n - 1 - t : t;
      double s = b[r];
      for(int k = rp[r]; k < rp[r + 1]; ++k) s -= val[k] * x[col[k]];
      x[r] = x[r] + omega * s / d[r];
    }
  }
  free(d);
}

static void sor_level(const olevel* L, const double* b, double* x, double omega, int flag) {
  sor_csr(L->ndof, L->rp, L->col, L->val, b, x, omega, flag);
}

/* one multigrid smoothing step: JacobiSmoother::smooth or SORSmoother::smooth
   (SORSmoother.C:29-33: sweep flag 0 for "sor", 2 for "ssor") */
static void smooth_step(const osolver* s, const olevel* L, const double* b, double* x) {
  if(s->smoother == 0) jacobi_smooth(L, b, x, s->jacobi_damp);
  else sor_level(L, b, x, s->sor_omega, s->smoother == 1 ? 0 : 2);
}

/* Cycle directions of CycleControl (CycleControl.C:17-111): a dummy first
   entry, then for each visit of a level with recursion count r: r times
   "down" followed by the coarser level's directions, then "up" ("end" on the
   finest level); the coarsest level contributes one "up". Recursion counts
   are 1 (V) or 2 (W, F) except 1 on the finest level; an F cycle resets a
   level's count to 1 after its first visit. */
enum { MG_DOWN = 0, MG_UP = 1, MG_END = 2, MG_DUMMY = 3 };

static void cycle_fill(int type, int* rec, int level, int finest, int* dirs, int* n) {
  if(level == 0) {
    dirs[(*n)++] = MG_UP;
    return;
  }
  for(int j = 0; j < rec[level]; ++j) {
    dirs[(*n)++] = MG_DOWN;
    cycle_fill(type, rec, level - 1, finest, dirs, n);
  }
  if(type == 2) rec[level] = 1;
  dirs[(*n)++] = level == finest ? MG_END : MG_UP;
}

/* directions into a malloc'ed array; returns its length */
static int cycle_dirs(int type, int nlev, int** out) {
  int* rec = malloc(sizeof(int) * nlev);
  for(int l = 0; l < nlev; ++l) rec[l] = type == 0 ? 1 : 2;
  rec[nlev - 1] = 1;
  /* at most 2^nlev visits per level */
  int cap = 4;
  for(int l = 0; l < nlev; ++l) cap *= 2;
  int* dirs = malloc(sizeof(int) * (size_t)cap * (nlev + 1));
  int n = 0;
  dirs[n++] = MG_DUMMY;
  cycle_fill(type, rec, nlev - 1, nlev - 1, dirs, &n);
  free(rec);
  *out = dirs;
  return n;
}

int orc_cycle_dirs(int type, int nlev, int* out, int cap) {
  int* d;
  const int n = cycle_dirs(type, nlev, &d);
  for(int i = 0; i < n && i < cap; ++i) out[i] = d[i];
  free(d);
  return n;
}

/* prolongate level l-1's iterate and add it, damped, to level l's active rows
   (Multigrid::update_solution_in_finer_grid, Multigrid.C:357-410) */
static void mg_prolong_add(orc_problem* P, const osolver* s, int l) {
  olevel* L = &P->lev[l];
  olevel* C = &P->lev[l - 1];
  prolongate(P, C, L, C->mg_sol, L->mg_aux);
  for(int j = 0; j < L->nactive; ++j) L->mg_sol[j] += s->mg_damp * L->mg_aux[j];
}

/* the coarsest level's loop with an iterative coarse smoother (Multigrid.C:
   200-240): the residual starts at 1e10, so at least one step runs; then
   smooth and recompute the defect norm until it is at most coarse_residual
   or coarse_max_iter steps are done */
static void coarse_iterate(const osolver* s, olevel* C) {
  osolver cs = *s;
  cs.smoother = s->coarse_smoother - 1;
  double res = 1e10;
  for(int i = 0; i < s->coarse_max_iter && res > s->coarse_residual; ++i) {
    smooth_step(&cs, C, C->mg_rhs, C->mg_sol);
    memcpy(C->mg_def, C->mg_rhs, sizeof(double) * C->ndof);
    apply_scaled_add(C, C->mg_sol, C->mg_def, -1.0);
    res = sqrt(ddot(C->ndof, C->mg_def, C->mg_def));
  }
}

/* Multigrid::cycle / cycle_step (Multigrid.C:157-308) with V, W or F control */
static void mg_cycle(orc_problem* P, const osolver* s, const double* v, double* z) {
  const int Lf = P->nlev - 1;
  olevel* F = &P->lev[Lf];
  memcpy(F->mg_rhs, v, sizeof(double) * F->ndof);
  memset(F->mg_sol, 0, sizeof(double) * F->nactive);
  memcpy(F->mg_sol + F->nactive, v + F->nactive, sizeof(double) * (F->ndof - F->nactive));
  if(P->nlev == 1) {
    for(int k = 0; k < s->n_pre + s->n_post; ++k) smooth_step(s, F, F->mg_rhs, F->mg_sol);
    memcpy(z, F->mg_sol, sizeof(double) * F->ndof);
    return;
  }
  int* dirs;
  const int nd = cycle_dirs(s->cycle, P->nlev, &dirs);
  int level = Lf;
  for(int i = 1; i < nd; ++i) {
    if(level == 0) {
      /* coarsest level: the direct solve or the coarse smoother loop, then up */
      olevel* C = &P->lev[0];
      if(s->coarse_smoother == 0) dense_solve(C, C->mg_rhs, C->mg_sol);
      else coarse_iterate(s, C);
      mg_prolong_add(P, s, 1);
      level = 1;
      continue;
    }
    olevel* L = &P->lev[level];
    if(dirs[i - 1] == MG_UP)
      for(int k = 0; k < s->n_post; ++k) smooth_step(s, L, L->mg_rhs, L->mg_sol);
    if(dirs[i] == MG_DOWN) {
      olevel* C = &P->lev[level - 1];
      for(int k = 0; k < s->n_pre; ++k) smooth_step(s, L, L->mg_rhs, L->mg_sol);
      memcpy(L->mg_def, L->mg_rhs, sizeof(double) * L->ndof);
      apply_scaled_add(L, L->mg_sol, L->mg_def, -1.0);
      memset(C->mg_sol, 0, sizeof(double) * C->ndof);
      restrict_defect(P, C, L, C->mg_rhs, L->mg_def);
      for(int j = C->nactive; j < C->ndof; ++j) C->mg_rhs[j] = 0;
      level = level - 1;
    } else if(dirs[i] == MG_UP) {
      mg_prolong_add(P, s, level + 1);
      level = level + 1;
    } else {
      break;
    }
  }
  free(dirs);
  memcpy(z, F->mg_sol, sizeof(double) * F->ndof);
}

static void precond(orc_problem* P, const osolver* s, const double* v, double* z) {
  olevel* F = &P->lev[P->nlev - 1];
  if(s->precond == 0) memcpy(z, v, sizeof(double) * F->ndof);
  else if(s->precond == 1)
    for(int i = 0; i < F->ndof; ++i) z[i] = v[i] / F->diag[i];
  else if(s->precond == 3 || s->precond == 4) {
    /* Iteration_sor::apply (Iteration_sor.C:125-140): sweep(s) from zero */
    memset(z, 0, sizeof(double) * F->ndof);
    sor_level(F, v, z, s->sor_omega, s->precond == 3 ? 0 : 2);
  } else mg_cycle(P, s, v, z);
}

static void givens_gen(double dx, double dy, double* cs, double* sn) {
  if(dy == 0.0) { *cs = 1; *sn = 0; }
  else if(fabs(dy) > fabs(dx)) { double t = dx / dy; *sn = 1.0 / sqrt(1.0 + t * t); *cs = t * *sn; }
  else { double t = dy / dx; *cs = 1.0 / sqrt(1.0 + t * t); *sn = t * *cs; }
}
static void givens_app(double* dx, double* dy, double cs, double sn) {
  double t = cs * *dx + sn * *dy;
  *dy = -sn * *dx + cs * *dy;
  *dx = t;
}

static int solve_fgmres(orc_problem* P, const osolver* s, ostats* st) {
  olevel* F = &P->lev[P->nlev - 1];
  const int n = F->ndof, m = s->restart;
  oconv cv = {s, 0, 0, 0};
  double* x = F->sol;
  double* r = malloc(sizeof(double) * n);
  double** v = malloc(sizeof(double*) * (m + 1));
  double** z = malloc(sizeof(double*) * m);
  for(int i = 0; i <= m; ++i) v[i] = malloc(sizeof(double) * n);
  for(int i = 0; i < m; ++i) z[i] = malloc(sizeof(double) * n);
  double* H = calloc((size_t)(m + 1) * m, sizeof(double));
#define HH(i, j) H[(size_t)(j) * (m + 1) + (i)]
  double* sv = calloc(m + 1, sizeof(double));
  double* cs = calloc(m + 1, sizeof(double));
  double* sn = calloc(m + 1, sizeof(double));
  double* y = calloc(m + 1, sizeof(double));
  memcpy(r, F->rhs, sizeof(double) * n);
  apply_scaled_add(F, x, r, -1.0);
  double resid = sqrt(ddot(n, r, r)), beta = resid;
  st->initial_residual = resid;
  int done = converged(&cv, resid, 0);
  int j = 1;
  while(!done && j <= s->max_iter) {
    for(int k = 0; k < n; ++k) v[0][k] = r[k] * (1.0 / beta);
    memset(sv, 0, sizeof(double) * (m + 1));
    sv[0] = beta;
    int i;
    for(i = 0; i < m && j <= s->max_iter; ++i, ++j) {
      precond(P, s, v[i], z[i]);
      apply(F, z[i], r);
      for(int k = 0; k <= i; ++k) {
        HH(k, i) = ddot(n, r, v[k]);
        for(int t = 0; t < n; ++t) r[t] += -HH(k, i) * v[k][t];
      }
      HH(i + 1, i) = sqrt(ddot(n, r, r));
      if(HH(i + 1, i) == 0.0) { fprintf(stderr, "oracle: unhappy breakdown\n"); done = 1; break; }
      for(int t = 0; t < n; ++t) v[i + 1][t] = r[t] * (1.0 / HH(i + 1, i));
      for(int k = 0; k < i; ++k) givens_app(&HH(k, i), &HH(k + 1, i), cs[k], sn[k]);
      givens_gen(HH(i, i), HH(i + 1, i), &cs[i], &sn[i]);
      givens_app(&HH(i, i), &HH(i + 1, i), cs[i], sn[i]);
      givens_app(&sv[i], &sv[i + 1], cs[i], sn[i]);
      resid = fabs(sv[i + 1]);
      st->iterations = j;
      if(converged(&cv, resid, j)) {
        /* update with i */
        memcpy(y, sv, sizeof(double) * (m + 1));
        for(int a = i; a >= 0; --a) {
          y[a] /= HH(a, a);
          for(int b = a - 1; b >= 0; --b) y[b] -= HH(b, a) * y[a];
        }
        memset(r, 0, sizeof(double) * n);
        for(int b = 0; b <= i; ++b)
          for(int t = 0; t < n; ++t) r[t] += y[b] * z[b][t];
        for(int t = 0; t < n; ++t) x[t] += r[t];
        done = 2;
        ++j;
        break;
      }
    }
    if(done) break;
    memcpy(y, sv, sizeof(double) * (m + 1));
    for(int a = m - 1; a >= 0; --a) {
      y[a] /= HH(a, a);
      for(int b = a - 1; b >= 0; --b) y[b] -= HH(b, a) * y[a];
    }
    memset(r, 0, sizeof(double) * n);
    for(int b = 0; b < m; ++b)
      for(int t = 0; t < n; ++t) r[t] += y[b] * z[b][t];
    for(int t = 0; t < n; ++t) x[t] += r[t];
    memcpy(r, F->rhs, sizeof(double) * n);
    apply_scaled_add(F, x, r, -1.0);
    beta = sqrt(ddot(n, r, r));
    if(converged(&cv, resid, j)) break;
  }
#undef HH
  st->final_residual = resid;
  st->converged = resid < s->res_tol || resid < s->res_red * st->initial_residual;
  for(int i = 0; i <= m; ++i) free(v[i]);
  for(int i = 0; i < m; ++i) free(z[i]);
  free(v); free(z); free(H); free(sv); free(cs); free(sn); free(y); free(r);
  return cv.diverged ? 2 : 0;
}

static int solve_cg(orc_problem* P, const osolver* s, ostats* st) {
  olevel* F = &P->lev[P->nlev - 1];
  const int n = F->ndof;
  oconv cv = {s, 0, 0, 0};
  double* x = F->sol;
  double *p = calloc(n, sizeof(double)), *zz = calloc(n, sizeof(double)), *qv = calloc(n, sizeof(double)),
         *r = calloc(n, sizeof(double));
  memcpy(r, F->rhs, sizeof(double) * n);
  apply_scaled_add(F, x, r, -1.0);
  double normb = sqrt(ddot(n, F->rhs, F->rhs));
  if(normb == 0.0) { normb = 1.0; memset(r, 0, sizeof(double) * n); memset(x, 0, sizeof(double) * n); }
  double resid = sqrt(ddot(n, r, r)) / normb, rho_1 = 0;
  st->initial_residual = resid;
  if(!converged(&cv, resid, 0)) {
    for(int i = 1; i <= s->max_iter; ++i) {
      precond(P, s, r, zz);
      double rho = ddot(n, r, zz);
      if(i == 1) memcpy(p, zz, sizeof(double) * n);
      else {
        double beta = rho / rho_1;
        for(int t = 0; t < n; ++t) p[t] *= beta;
        for(int t = 0; t < n; ++t) p[t] += zz[t];
      }
      apply(F, p, qv);
      double alpha = rho / ddot(n, p, qv);
      for(int t = 0; t < n; ++t) x[t] += alpha * p[t];
      for(int t = 0; t < n; ++t) r[t] += -alpha * qv[t];
      resid = sqrt(ddot(n, r, r)) / normb;
      st->iterations = i;
      if(converged(&cv, resid, i)) break;
      rho_1 = rho;
    }
  }
  st->final_residual = resid;
  st->converged = resid < s->res_tol || resid < s->res_red * st->initial_residual;
  free(p); free(zz); free(qv); free(r);
  return cv.diverged ? 2 : 0;
}

/* Iteration_sor::iterate (Iteration_sor.C:27-120) */
static int solve_sor(orc_problem* P, const osolver* s, ostats* st, int flag) {
  olevel* F = &P->lev[P->nlev - 1];
  const int n = F->ndof;
  oconv cv = {s, 0, 0, 0};
  double* x = F->sol;
  double* r = malloc(sizeof(double) * n);
  memcpy(r, F->rhs, sizeof(double) * n);
  apply_scaled_add(F, x, r, -1.0);
  double normb = sqrt(ddot(n, F->rhs, F->rhs));
  if(normb == 0.0) { normb = 1.0; memset(r, 0, sizeof(double) * n); memset(x, 0, sizeof(double) * n); }
  double resid = sqrt(ddot(n, r, r)) / normb;
  st->initial_residual = resid;
  if(!converged(&cv, resid, 0)) {
    for(int i = 1; i <= s->max_iter; ++i) {
      sor_level(F, F->rhs, x, s->sor_omega, flag);
      memcpy(r, F->rhs, sizeof(double) * n);
      apply_scaled_add(F, x, r, -1.0);
      resid = sqrt(ddot(n, r, r)) / normb;
      st->iterations = i;
      if(converged(&cv, resid, i)) break;
    }
  }
  st->final_residual = resid;
  st->converged = resid < s->res_tol || resid < s->res_red * st->initial_residual;
  free(r);
  return cv.diverged ? 2 : 0;
}

/* ----------------------------------------------------------------------- */
/* exported API                                                             */
/* ----------------------------------------------------------------------- */
orc_problem* orc_create(int geometry, int structured_n, int n_refine, int order, int supg, int linear_quad,
                        int mg_levels, int example, double eps, const double* b, double c, double f0) {
  orc_problem* P = calloc(1, sizeof(orc_problem));
  omesh* meshes = calloc(n_refine + 1, sizeof(omesh));
  if(geometry == 10 || geometry == 11) structured_mesh(structured_n, geometry == 11, &meshes[0]);
  else default_mesh(geometry, &meshes[0]);
  for(int i = 0; i < n_refine; ++i) refine(&meshes[i], &meshes[i + 1]);
  int keep = mg_levels < 1 ? 1 : mg_levels;
  if(keep > n_refine + 1) keep = n_refine + 1;
  P->nlev = keep;
  P->lev = calloc(keep, sizeof(olevel));
  const int shape = meshes[0].shape;
  P->el = make_elem(shape, order);
  P->qd = make_quad(shape, (linear_quad ? 2 : 3) * order);
  P->ex.kind = example;
  P->ex.t = 0;
  P->ex.eps = eps;
  for(int k = 0; k < 3; ++k) P->ex.b[k] = b ? b[k] : 0.0;
  P->ex.c = c;
  P->ex.f0 = f0;
  P->supg = supg;
  for(int i = 0; i <= n_refine; ++i) {
    int l = i - (n_refine + 1 - keep);
    if(l < 0) { mesh_free(&meshes[i]); continue; }
    olevel* L = &P->lev[l];
    L->m = meshes[i];
    build_neighbours(L);
    build_space(P, L);
    build_structure(L);
    L->mg_sol = calloc(L->ndof, sizeof(double));
    L->mg_rhs = calloc(L->ndof, sizeof(double));
    L->mg_def = calloc(L->ndof, sizeof(double));
    L->mg_aux = calloc(L->ndof, sizeof(double));
  }
  free(meshes);
  return P;
}

void orc_destroy(orc_problem* P);

/* a problem on the leaves of `adapt_steps` indicator refinements
   (TDomain::RefineByIndicator without conforming closure) of the coarse
   mesh refined `n_refine` times: one level, 2D, order 1 or 2, the conforming
   closure on triangles only (NULL otherwise); indicator(x, ctx) is
   evaluated at the cells' vertices */
orc_problem* orc_create_adaptive(int geometry, int n_refine, int adapt_steps,
                                 double (*indicator)(const double*, void*), void* ctx, int conforming, int order,
                                 int supg, int linear_quad, int example, double eps, const double* b, double c,
                                 double f0) {
  omesh cur;
  default_mesh(geometry, &cur);
  if((order != 1 && order != 2) || DIM[cur.shape] != 2 || (conforming && cur.shape != TRI)) {
    mesh_free(&cur);
    return NULL;
  }
  for(int i = 0; i < n_refine; ++i) {
    omesh next;
    refine(&cur, &next);
    mesh_free(&cur);
    cur = next;
  }
  oindicator ind = {indicator, ctx};
  omesh leaves;
  vdict mids;
  adapt_mesh(&cur, &ind, adapt_steps, &leaves, &mids, conforming);
  orc_problem* P = calloc(1, sizeof(orc_problem));
  P->nlev = 1;
  P->lev = calloc(1, sizeof(olevel));
  P->el = make_elem(leaves.shape, order);
  P->qd = make_quad(leaves.shape, (linear_quad ? 2 : 3) * order);
  P->ex.kind = example;
  P->ex.eps = eps;
  for(int k = 0; k < 3; ++k) P->ex.b[k] = b ? b[k] : 0.0;
  P->ex.c = c;
  P->ex.f0 = f0;
  P->supg = supg;
  olevel* L = &P->lev[0];
  L->m = leaves;
  build_neighbours(L);
  const int rc = build_space_hanging(P, L, &mids);
  free(mids.slot);
  free(mids.key);
  if(rc) { orc_destroy(P); return NULL; }
  build_structure(L);
  hanging_structure(L);
  L->mg_sol = calloc(L->ndof, sizeof(double));
  L->mg_rhs = calloc(L->ndof, sizeof(double));
  L->mg_def = calloc(L->ndof, sizeof(double));
  L->mg_aux = calloc(L->ndof, sizeof(double));
  return P;
}

/* hanging DOFs in the sorted order: out[4t] = hanging DOF, out[4t+1..3] =
   its coupling DOFs (-1 beyond the count), w[3t..3t+2] their weights (0
   beyond); returns the count */
int orc_hanging(orc_problem* P, int* out, double* w) {
  olevel* L = &P->lev[P->nlev - 1];
  for(int t = 0; t < L->nhang; ++t) {
    const int h = L->hang_order[t];
    if(out) {
      out[4 * t] = L->nactive_nh + h;
      for(int k = 0; k < MAXCPL; ++k) out[4 * t + 1 + k] = k < L->hang_ncpl[h] ? L->hang_cpl[MAXCPL * h + k] : -1;
    }
    if(w)
      for(int k = 0; k < MAXCPL; ++k) w[3 * t + k] = k < L->hang_ncpl[h] ? L->hang_w[MAXCPL * h + k] : 0.0;
  }
  return L->nhang;
}

void orc_destroy(orc_problem* P) {
  if(!P) return;
  for(int l = 0; l < P->nlev; ++l) {
    olevel* L = &P->lev[l];
    mesh_free(&L->m);
    free(L->dof); free(L->rp); free(L->col); free(L->val); free(L->rhs); free(L->sol);
    free(L->mg_sol); free(L->mg_rhs); free(L->mg_def); free(L->mg_aux); free(L->lu); free(L->piv);
    free(L->nb_cell); free(L->nb_joint); free(L->comp); free(L->qq); free(L->diag);
    free(L->aval); free(L->mval); free(L->hang_cpl); free(L->hang_order); free(L->hang_ncpl); free(L->hang_w);
  }
  free(P->u_old); free(P->u_old2); free(P->f_old);
  free(P->lev);
  free(P);
}

static olevel* lvl(orc_problem* P, int l) { return &P->lev[l < 0 ? P->nlev + l : l]; }

void orc_info(orc_problem* P, int level, long long* info) {
  olevel* L = lvl(P, level);
  info[0] = L->m.ncells; info[1] = L->ndof; info[2] = L->nactive; info[3] = L->nnz;
  info[4] = L->nl; info[5] = P->qd.n; info[6] = P->nlev; info[7] = L->ndir;
}
void orc_dofmap(orc_problem* P, int level, int* out) {
  olevel* L = lvl(P, level);
  memcpy(out, L->dof, sizeof(int) * L->m.ncells * L->nl);
}
void orc_pattern(orc_problem* P, int level, int* rp, int* col) {
  olevel* L = lvl(P, level);
  memcpy(rp, L->rp, sizeof(int) * (L->ndof + 1));
  memcpy(col, L->col, sizeof(int) * L->nnz);
}
void orc_cell_coords(orc_problem* P, int level, double* out) {
  olevel* L = lvl(P, level);
  const int nv = NV[L->m.shape];
  for(int c = 0; c < L->m.ncells; ++c)
    for(int a = 0; a < nv; ++a) memcpy(&out[(c * nv + a) * 3], &L->m.X[3 * L->m.cv[c * nv + a]], 24);
}
static void extract_diag(olevel* L) {
  if(!L->diag) L->diag = malloc(sizeof(double) * L->ndof);
  for(int r = 0; r < L->ndof; ++r) L->diag[r] = diag_of(L, r);
}

void orc_assemble(orc_problem* P) {
  for(int l = 0; l < P->nlev; ++l) {
    assemble_level(P, &P->lev[l], 0, P->lev[l].m.ncells);
    if(P->lev[l].nhang && orc_fault != 9) apply_hanging(&P->lev[l]);
    extract_diag(&P->lev[l]);
  }
  olevel* F = lvl(P, -1);
  memset(F->sol, 0, sizeof(double) * F->nactive);
  memcpy(F->sol + F->nactive, F->rhs + F->nactive, sizeof(double) * (F->ndof - F->nactive));
}
void orc_matrix(orc_problem* P, int level, double* v) { olevel* L = lvl(P, level); memcpy(v, L->val, sizeof(double) * L->nnz); }
void orc_rhs(orc_problem* P, int level, double* v) { olevel* L = lvl(P, level); memcpy(v, L->rhs, sizeof(double) * L->ndof); }
void orc_solution(orc_problem* P, double* v) { olevel* L = lvl(P, -1); memcpy(v, L->sol, sizeof(double) * L->ndof); }
void orc_set_solution(orc_problem* P, const double* v) { olevel* L = lvl(P, -1); memcpy(L->sol, v, sizeof(double) * L->ndof); }
void orc_apply(orc_problem* P, int level, const double* x, double* y, double a, int mode) {
  olevel* L = lvl(P, level);
  if(mode == 0) apply(L, x, y);
  else apply_scaled_add(L, x, y, a);
}

int orc_solve(orc_problem* P, const osolver* s, ostats* st) {
  memset(st, 0, sizeof(*st));
  if(s->method == 0 || (s->precond == 2 && P->nlev > 1)) dense_factor(&P->lev[0]);
  if(s->method == 0) {
    olevel* F = lvl(P, -1);
    dense_solve(F, F->rhs, F->sol);
    st->converged = 1;
    return 0;
  }
  if(s->method == 1) return solve_cg(P, s, st);
  if(s->method == 4 || s->method == 5) return solve_sor(P, s, st, s->method == 4 ? 0 : 2);
  return solve_fgmres(P, s, st);
}

/* ---- time-dependent problems ----
   TimeConvectionDiffusion<d>::assemble_initial_time / assemble / solve
   (src/System/TimeConvectionDiffusion.C:259-452) with the schemes of
   TimeDiscretization (src/TimeDiscRoutines/TimeDiscretizations.C:117-284,
   426-483): every step assembles the stiffness matrix A, the mass matrix M
   and the rhs f at the new time on every level; backward Euler solves
   (M + tau A) u = tau f + M u_old, Crank-Nicolson (M + tau/2 A) u =
   tau/2 (f + f_old) + M u_old - tau/2 A u_old, BDF2 (M + 2/3 tau A) u =
   2/3 tau f + 4/3 M u_old - 1/3 M u_old2 with its first step by backward
   Euler or Crank-Nicolson. Only active rows take the scheme; Dirichlet rows
   keep A's identity rows and the boundary values at the new time. */
static const double BDF2[3] = {4. / 3., -1. / 3., 2. / 3.};

/* y += a * V x on active rows for a value array V of the level's pattern */
static void csr_add_active(const olevel* L, const double* V, const double* x, double* y, double a) {
  for(int i = 0; i < L->nactive; ++i) {
    double v = 0;
    for(int j = L->rp[i]; j < L->rp[i + 1]; ++j) v += V[j] * x[L->col[j]];
    y[i] += a * v;
  }
}

static void tcd_assemble_all(orc_problem* P) {
  for(int l = 0; l < P->nlev; ++l) assemble_level(P, &P->lev[l], 0, P->lev[l].m.ncells);
}

/* time parameters and matrix storage only, no assembly (orc_assemble_cells
   then assembles the time-dependent forms of a range of finest cells) */
int orc_tcd_config(orc_problem* P, int scheme, int bdf_first, double tau, double t0) {
  if(scheme < 0 || scheme > 2 || (bdf_first != 0 && bdf_first != 1) || !(tau > 0)) return 1;
  P->tcd = 1;
  P->scheme = scheme;
  P->bdf_first = bdf_first;
  P->tau = tau;
  P->time = t0;
  P->step = 0;
  P->ex.t = t0;
  for(int l = 0; l < P->nlev; ++l) {
    olevel* L = &P->lev[l];
    if(!L->aval) {
      L->aval = calloc(L->nnz, sizeof(double));
      L->mval = calloc(L->nnz, sizeof(double));
    }
  }
  return 0;
}

int orc_tcd_init(orc_problem* P, int scheme, int bdf_first, double tau, double t0) {
  if(orc_tcd_config(P, scheme, bdf_first, tau, t0)) return 1;
  tcd_assemble_all(P);
  olevel* F = lvl(P, -1);
  const int n = F->ndof, nl = F->nl;
  /* initial condition: the exact solution at the nodal points */
  memset(F->sol, 0, sizeof(double) * n);
  for(int c = 0; c < F->m.ncells; ++c)
    for(int i = 0; i < nl; ++i) {
      double x[3], e[5];
      node_xyz(P, &F->m, c, i, x);
      if(ex_exact(&P->ex, x[0], x[1], P->el.dim == 3 ? x[2] : 0.0, e)) F->sol[F->dof[c * nl + i]] = e[0];
    }
  free(P->u_old); free(P->u_old2); free(P->f_old);
  P->u_old = malloc(sizeof(double) * n);
  P->u_old2 = malloc(sizeof(double) * n);
  P->f_old = malloc(sizeof(double) * n);
  memcpy(P->u_old, F->sol, sizeof(double) * n);
  memcpy(P->u_old2, F->sol, sizeof(double) * n);
  memcpy(P->f_old, F->rhs, sizeof(double) * n);
  return 0;
}

int orc_tcd_step(orc_problem* P, const osolver* s, ostats* st, double* time) {
  if(!P->tcd) return 1;
  P->step += 1;
  const int eff = P->scheme == 2 ? (P->step <= 1 ? P->bdf_first : 2) : P->scheme;
  const double tau = P->tau;
  P->time += tau;
  P->ex.t = P->time;
  tcd_assemble_all(P);
  olevel* F = lvl(P, -1);
  const int n = F->ndof, na = F->nactive;
  double* f_new = malloc(sizeof(double) * n);
  memcpy(f_new, F->rhs, sizeof(double) * n);
  double* r = F->rhs;
  if(eff == 2) {
    for(int i = 0; i < na; ++i) r[i] *= BDF2[2] * tau;
    csr_add_active(F, F->mval, P->u_old, r, BDF2[0]);
    csr_add_active(F, F->mval, P->u_old2, r, BDF2[1]);
  } else if(eff == 0) {
    for(int i = 0; i < na; ++i) r[i] *= tau;
    csr_add_active(F, F->mval, P->u_old, r, 1.0);
  } else {
    for(int i = 0; i < na; ++i) r[i] = 0.5 * tau * r[i] + 0.5 * tau * P->f_old[i];
    csr_add_active(F, F->mval, P->u_old, r, 1.0);
    csr_add_active(F, F->aval, P->u_old, r, -0.5 * tau);
  }
  memcpy(r + na, f_new + na, sizeof(double) * (n - na));
  const double factor = eff == 0 ? tau : (eff == 1 ? 0.5 * tau : BDF2[2] * tau);
  for(int l = 0; l < P->nlev; ++l) {
    olevel* L = &P->lev[l];
    for(int i = 0; i < L->ndof; ++i)
      for(int j = L->rp[i]; j < L->rp[i + 1]; ++j)
        L->val[j] = i < L->nactive ? factor * L->aval[j] + L->mval[j] : L->aval[j];
    extract_diag(L);
    if(L->lu) { free(L->lu); L->lu = NULL; free(L->piv); L->piv = NULL; }
  }
  memcpy(F->sol + na, r + na, sizeof(double) * (n - na));
  memcpy(P->f_old, f_new, sizeof(double) * n);
  free(f_new);
  const int rc = orc_solve(P, s, st);
  memcpy(P->u_old2, P->u_old, sizeof(double) * n);
  memcpy(P->u_old, F->sol, sizeof(double) * n);
  if(time) *time = P->time;
  return rc;
}

/* ---- algebraic flux correction: Kuzmin / BJK17 limiter, fixed point rhs /
   fixed point matrix ----
   ConvectionDiffusion_AFC<d>::assemble / solve / dynamic_damping
   (src/System/ConvectionDiffusion_AFC.C:132-600, 1032-1180) and
   AlgebraicFluxCorrection::steady_state_algorithm (non-MPI path,
   src/AssembleRoutines/AlgebraicFluxCorrection.C:2058-2745),
   compute_artificial_diffusion_matrix (:82-165), AFC_Compute_New_Iterate
   (:3042-3080), correct_dirichlet_hanging_rows (:2992-3015). The Galerkin
   matrix A is assembled with its Dirichlet rows (assemble_dirichlet_rows);
   D = artificial diffusion, E = A + D, the solver matrix S = E with identity
   Dirichlet rows; the scheme iterates S u = f + fbar(u) with the limited
   fluxes fbar and the reference's dynamic damping. */
static int transposed_pos(const olevel* L, int c, int i) {
  int e = L->rp[c];
  for(; e < L->rp[c + 1]; ++e)
    if(L->col[e] == i) break;
  return e;
}

static void afc_diffusion(const olevel* L, const double* A, double* D) {
  memset(D, 0, sizeof(double) * L->nnz);
  for(int i = 0; i < L->ndof; ++i)
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int j = L->col[l];
      if(j == i) continue;
      const double kij = A[l];
      double kji = 0;
      for(int ll = L->rp[j]; ll < L->rp[j + 1]; ++ll)
        if(L->col[ll] == i) { kji = A[ll]; break; }
      double m = -kij; /* std::min({-kij, 0.0, -kji}) keeps the first smallest */
      if(0.0 < m) m = 0.0;
      if(-kji < m) m = -kji;
      D[l] = m;
    }
  for(int i = 0; i < L->ndof; ++i) {
    double val = 0;
    int ll = -1;
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      val += D[l];
      if(L->col[l] == i) ll = l;
    }
    D[ll] = -val;
  }
}

typedef struct {
  double *A, *D, *E, *f, *F, *alpha, *Pp, *Pm, *Qp, *Qm, *Rp, *Rm, *gamma;
  double *umax, *umin, *q, *df; /* BJK17 bounds and q_i; the Newton matrix */
} oafc;

/* Compute_Parameter_For_Linearity_Preservation (AlgebraicFluxCorrection.C:
   1295-1470), P1 triangles and tetrahedra: per vertex the largest distance
   to another vertex of a cell over the smallest height of the vertex over
   the opposite side (2D: the distance to a neighbour when the angle at it is
   not acute) */
static void afc_gamma(const orc_problem* P, const olevel* L, double* gamma) {
  const omesh* m = &L->m;
  const int n = L->ndof, nv = NV[m->shape], d = P->el.dim;
  double* par = malloc(sizeof(double) * 2 * n);
  for(int i = 0; i < n; ++i) { par[i] = 0.0; par[n + i] = 4711.0; }
  for(int c = 0; c < m->ncells; ++c) {
    double x[4], y[4], z[4];
    for(int a = 0; a < nv; ++a) {
      const double* X = &m->X[3 * m->cv[c * nv + a]];
      x[a] = X[0]; y[a] = X[1]; z[a] = X[2];
    }
    ogeo g;
    geo(m, c, &g);
    const double area = fabs(g.det) / (d == 2 ? 2.0 : 6.0);
    for(int j = 0; j < nv; ++j) {
      const int index = L->dof[c * L->nl + j];
      double dist, dist_max;
      const int j1 = (j + 1) % nv, j2 = (j + 2) % nv, j3 = (j + 3) % nv;
      if(d == 2) {
        if((x[j2] - x[j1]) * (x[j] - x[j1]) + (y[j2] - y[j1]) * (y[j] - y[j1]) <= 0) {
          dist = sqrt((x[j] - x[j1]) * (x[j] - x[j1]) + (y[j] - y[j1]) * (y[j] - y[j1]));
          dist_max = sqrt((x[j] - x[j2]) * (x[j] - x[j2]) + (y[j] - y[j2]) * (y[j] - y[j2]));
        } else if((x[j1] - x[j2]) * (x[j] - x[j2]) + (y[j1] - y[j2]) * (y[j] - y[j2]) <= 0) {
          dist = sqrt((x[j] - x[j2]) * (x[j] - x[j2]) + (y[j] - y[j2]) * (y[j] - y[j2]));
          dist_max = sqrt((x[j] - x[j1]) * (x[j] - x[j1]) + (y[j] - y[j1]) * (y[j] - y[j1]));
        } else {
          double edge = (x[j1] - x[j2]) * (x[j1] - x[j2]);
          edge += (y[j1] - y[j2]) * (y[j1] - y[j2]);
          edge = sqrt(edge);
          dist = 2. * area / edge;
          dist_max = sqrt((x[j] - x[j1]) * (x[j] - x[j1]) + (y[j] - y[j1]) * (y[j] - y[j1]));
          const double dm2 = sqrt((x[j] - x[j2]) * (x[j] - x[j2]) + (y[j] - y[j2]) * (y[j] - y[j2]));
          if(dm2 > dist_max) dist_max = dm2;
        }
      } else {
        const double d1 = sqrt((x[j] - x[j1]) * (x[j] - x[j1]) + (y[j] - y[j1]) * (y[j] - y[j1]) +
                               (z[j] - z[j1]) * (z[j] - z[j1]));
        const double d2 = sqrt((x[j] - x[j2]) * (x[j] - x[j2]) + (y[j] - y[j2]) * (y[j] - y[j2]) +
                               (z[j] - z[j2]) * (z[j] - z[j2]));
        const double d3 = sqrt((x[j] - x[j3]) * (x[j] - x[j3]) + (y[j] - y[j3]) * (y[j] - y[j3]) +
                               (z[j] - z[j3]) * (z[j] - z[j3]));
        dist_max = fmax(fmax(d1, d2), d3);
        const double A = (z[j3] - z[j1]) * (y[j2] - y[j1]) - (z[j2] - z[j1]) * (y[j3] - y[j1]);
        const double B = (z[j2] - z[j1]) * (x[j3] - x[j1]) - (z[j3] - z[j1]) * (x[j2] - x[j1]);
        const double C = (y[j3] - y[j1]) * (x[j2] - x[j1]) - (x[j3] - x[j1]) * (y[j2] - y[j1]);
        dist = (6.0 * area) / sqrt(A * A + B * B + C * C);
      }
      if(dist_max > par[index]) par[index] = dist_max;
      if(dist < par[index + n]) par[index + n] = dist;
    }
  }
  for(int i = 0; i < n; ++i) gamma[i] = par[i] / par[i + n];
  free(par);
}

/* BJK17 limiter (steady_state_algorithm, :2200-2270, 2527-2560): local
   bounds from the row's neighbours, Q from gamma and D, alpha the smaller
   of the two rows' factors; rhs_i += sum_j alpha_ij F_ij over the row */
static void afc_bjk_rhs(const olevel* L, oafc* W, const double* u, double* rhs) {
  const int n = L->ndof, na = L->nactive;
  for(int i = 0; i < n; ++i) {
    double umin = u[i], umax = u[i], q = 0;
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      if(u[c] < umin) umin = u[c];
      if(u[c] > umax) umax = u[c];
      if(c != i) q += W->D[l];
    }
    q *= W->gamma[i];
    W->umin[i] = umin;
    W->umax[i] = umax;
    W->q[i] = q;
    double Pp = 0, Pm = 0;
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      W->F[l] = W->D[l] * (u[c] - u[i]);
      if(c == i) continue;
      if(W->F[l] > 0) Pp += W->F[l];
      if(W->F[l] < 0) Pm += W->F[l];
    }
    const double Qp = q * (u[i] - umax), Qm = q * (u[i] - umin);
    W->Pp[i] = Pp; W->Pm[i] = Pm; W->Qp[i] = Qp; W->Qm[i] = Qm;
    W->Rp[i] = W->Rm[i] = 1.0;
    if(fabs(Pp) > 0) { W->Rp[i] = Qp / Pp; if(W->Rp[i] > 1) W->Rp[i] = 1; }
    if(fabs(Pm) > 0) { W->Rm[i] = Qm / Pm; if(W->Rm[i] > 1) W->Rm[i] = 1; }
  }
  for(int i = na; i < n; ++i) W->Rp[i] = W->Rm[i] = 1.0;
  memcpy(rhs, W->f, sizeof(double) * n);
  for(int i = 0; i < n; ++i) {
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      double a = 1;
      if(W->F[l] > 0) { a = W->Rp[i]; if(W->Rm[c] < a) a = W->Rm[c]; }
      if(W->F[l] < 0) { a = W->Rm[i]; if(W->Rp[c] < a) a = W->Rp[c]; }
      W->alpha[l] = c == i ? 1.0 : a;
    }
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) rhs[i] += W->alpha[l] * W->F[l];
  }
  memcpy(rhs + na, u + na, sizeof(double) * (n - na));
}

/* rhs = f + limited fluxes of u (Kuzmin limiter); Dirichlet rows = u */
static void afc_kuzmin_rhs(const olevel* L, oafc* W, const double* u, double* rhs) {
  const int n = L->ndof, na = L->nactive;
  const double* E = W->E;
  for(int i = 0; i < n; ++i)
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) W->F[l] = W->D[l] * (u[L->col[l]] - u[i]);
  for(int i = 0; i < n; ++i) W->Pp[i] = W->Pm[i] = W->Qp[i] = W->Qm[i] = 0;
  for(int i = 0; i < n; ++i)
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      if(c == i) continue;
      const int jj = transposed_pos(L, c, i);
      if(E[jj] > E[l]) continue;
      const double F = W->F[l];
      if(F > 0) { W->Pp[i] += F; W->Qp[c] += F; W->Qm[i] -= F; }
      if(F < 0) { W->Pm[i] += F; W->Qp[i] -= F; W->Qm[c] += F; }
    }
  for(int i = 0; i < n; ++i) {
    W->Rp[i] = W->Rm[i] = 1.0;
    if(fabs(W->Pp[i]) > 0) { W->Rp[i] = W->Qp[i] / W->Pp[i]; if(W->Rp[i] > 1) W->Rp[i] = 1; }
    if(fabs(W->Pm[i]) > 0) { W->Rm[i] = W->Qm[i] / W->Pm[i]; if(W->Rm[i] > 1) W->Rm[i] = 1; }
  }
  for(int i = na; i < n; ++i) W->Rp[i] = W->Rm[i] = 1.0;
  for(int i = 0; i < n; ++i) {
    int diag = i;
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      if(c == i) diag = l;
      const int jj = transposed_pos(L, c, i);
      if(E[jj] > E[l]) continue;
      double a = 1;
      if(W->F[l] > 0) a = W->Rp[i];
      if(W->F[l] < 0) a = W->Rm[i];
      W->alpha[l] = a;
      W->alpha[jj] = a;
    }
    W->alpha[diag] = 1.0;
  }
  memcpy(rhs, W->f, sizeof(double) * n);
  for(int i = 0; i < n; ++i)
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      if(c == i) continue;
      const int jj = transposed_pos(L, c, i);
      if(E[jj] > E[l]) continue;
      rhs[i] += W->alpha[l] * W->F[l];
      rhs[c] -= W->alpha[l] * W->F[l];
    }
  memcpy(rhs + na, u + na, sizeof(double) * (n - na));
}

/* fixed point matrix scheme (steady_state_algorithm :2746-2790): with the
   limiter's alpha, S = E - tau0 alpha D off the diagonal, S_ii = E_ii +
   tau0 sum_j alpha_ij d_ij; rhs = f + tau1 (sum_j alpha_ij d_ij u_j -
   (sum_j alpha_ij d_ij) u_i) in entry order; Dirichlet rows identity / u */
static void afc_fpm(olevel* L, const oafc* W, const double* u, double* rhs, double tau0) {
  const int n = L->ndof, na = L->nactive;
  const double tau1 = 1.0 - tau0;
  for(int i = 0; i < n; ++i) {
    const int b = L->rp[i], e = L->rp[i + 1];
    if(i >= na) {
      for(int l = b; l < e; ++l) L->val[l] = L->col[l] == i ? 1.0 : 0.0;
      rhs[i] = u[i];
      continue;
    }
    double r = W->f[i];
    for(int l = b; l < e; ++l) {
      const int c = L->col[l];
      double v = W->E[l];
      if(c != i) {
        v -= W->alpha[l] * W->D[l] * tau0;
        r += W->alpha[l] * W->D[l] * u[c] * tau1;
      } else {
        double sum_flux = 0.0;
        for(int ll = b; ll < e; ++ll)
          if(L->col[ll] != i) sum_flux += W->alpha[ll] * W->D[ll];
        v += sum_flux * tau0;
        r -= sum_flux * u[i] * tau1;
      }
      L->val[l] = v;
    }
    rhs[i] = r;
  }
  extract_diag(L);
  if(L->lu) { free(L->lu); L->lu = NULL; free(L->piv); L->piv = NULL; }
}

/* sums of D over row r for the derivative of Q and P at a diagonal entry
   (Compute_Jacobian_times_flux_Kuzmin, :296-345): over the entries with
   F < 0 (sign -1) or F > 0 (sign +1); with `upwind`, only the entries whose
   E is at least the transposed entry's */
static double row_d_sum(const olevel* L, const oafc* W, int r, int sign, int upwind) {
  double s = 0;
  for(int m = L->rp[r]; m < L->rp[r + 1]; ++m) {
    const int ok = sign > 0 ? W->F[m] > 0 : W->F[m] < 0;
    if(!ok) continue;
    if(upwind && !(W->E[m] >= W->E[transposed_pos(L, L->col[m], r)])) continue;
    s += W->D[m];
  }
  return s;
}

/* Compute_Jacobian_times_flux_Kuzmin (AlgebraicFluxCorrection.C:224-578):
   the sum over the row of i of the derivative of the limiter factor of the
   link (i, k) with respect to u_j times the flux F_ik, with the reference's
   case distinctions. The reference keeps E_kj in an int (a_kj) in the
   downwind branch; the comparison below does the same. */
static double jac_kuzmin(const olevel* L, const oafc* W, int i, int j, int pij) {
  const double eps = 1e-14, epsinv = 1.0 / eps;
  const double *E = W->E, *F = W->F, *D = W->D;
  const int a_ji = transposed_pos(L, j, i);
  double sum = 0;
  for(int k = L->rp[i]; k < L->rp[i + 1]; ++k) {
    const int kk = L->col[k];
    int a_kj = -1, a_ki = L->rp[kk];
    for(int jj = L->rp[kk]; jj < L->rp[kk + 1]; ++jj) {
      if(L->col[jj] == j) a_kj = jj;
      if(L->col[jj] == i) a_ki = jj;
    }
    if(E[k] >= E[a_ki]) {
      if(F[k] > 0 && fabs(W->Rp[i] - 1) < eps) continue;
      if(F[k] < 0 && fabs(W->Rm[i] - 1) < eps) continue;
      if(fabs(F[k]) < eps) continue;
      if(i != j && fabs(D[pij]) < 1e-14) continue;
      if(F[k] > 0) {
        double DQ, DP;
        if(i != j) DQ = F[pij] >= 0 ? 0 : -D[pij];
        else DQ = row_d_sum(L, W, i, -1, 0);
        if(i != j) DP = F[pij] <= 0 ? 0 : (E[pij] >= E[a_ji] ? D[pij] : 0);
        else DP = -row_d_sum(L, W, i, +1, 1);
        const double val = DQ / W->Pp[i] - W->Qp[i] * DP / (W->Pp[i] * W->Pp[i]);
        if(fabs(val) < epsinv) sum += val * F[k];
        continue;
      }
      if(F[k] < 0) {
        double DQ, DP;
        if(i != j) DQ = F[pij] <= 0 ? 0 : -D[pij];
        else DQ = row_d_sum(L, W, i, +1, 0);
        if(i != j) DP = F[pij] >= 0 ? 0 : (E[pij] >= E[a_ji] ? D[pij] : 0);
        else DP = -row_d_sum(L, W, i, -1, 1);
        const double val = DQ / W->Pm[i] - W->Qm[i] * DP / (W->Pm[i] * W->Pm[i]);
        if(fabs(val) < epsinv) sum += val * F[k];
        continue;
      }
    } else {
      if(F[k] > 0 && fabs(W->Rm[kk] - 1) < eps) continue;
      if(F[k] < 0 && fabs(W->Rp[kk] - 1) < eps) continue;
      if(fabs(F[k]) < eps) continue;
      if(a_kj < 0) continue;
      const double f_kj = F[a_kj], d_kj = D[a_kj];
      const double e_kj = orc_fault == 5 ? E[a_kj] : (double)(int)E[a_kj];
      if(kk != j && fabs(d_kj) < 1e-14) continue;
      const int a_jk = transposed_pos(L, j, kk);
      if(F[k] > 0) {
        double DQ, DP;
        if(kk != j) DQ = f_kj <= 0 ? 0 : -d_kj;
        else DQ = row_d_sum(L, W, kk, +1, 0);
        if(kk != j) DP = f_kj >= 0 ? 0 : (e_kj >= E[a_jk] ? d_kj : 0);
        else DP = -row_d_sum(L, W, kk, -1, 1);
        const double val = DQ / W->Pm[kk] - W->Qm[kk] * DP / (W->Pm[kk] * W->Pm[kk]);
        if(fabs(val) < epsinv) sum += val * F[k];
        continue;
      }
      if(F[k] < 0) {
        double DQ, DP;
        if(kk != j) DQ = f_kj >= 0 ? 0 : -d_kj;
        else DQ = row_d_sum(L, W, kk, -1, 0);
        if(kk != j) DP = f_kj <= 0 ? 0 : (e_kj >= E[a_jk] ? d_kj : 0);
        else DP = -row_d_sum(L, W, kk, +1, 1);
        const double val = DQ / W->Pp[kk] - W->Qp[kk] * DP / (W->Pp[kk] * W->Pp[kk]);
        if(fabs(val) < epsinv) sum += val * F[k];
        continue;
      }
    }
  }
  return sum;
}

/* Compute_Jacobian_times_flux_BJK17 (AlgebraicFluxCorrection.C:957-1140) */
static double jac_bjk(const olevel* L, const oafc* W, double u_j, int i, int j, int pij) {
  const double eps = 1e-14, epsinv = 1.0 / eps;
  const double *F = W->F, *D = W->D;
  double sum = 0;
  for(int k = L->rp[i]; k < L->rp[i + 1]; ++k) {
    if(F[k] > 0 && fabs(W->Rp[i] - 1.0) < eps) continue;
    if(F[k] < 0 && fabs(W->Rm[i] - 1.0) < eps) continue;
    if(fabs(F[k]) < eps) continue;
    const int ck = L->col[k];
    double DQ = 0.0, DP = 0.0;
    if(F[k] > 0) {
      if(W->Rp[i] <= W->Rm[ck]) {
        if(i != j) {
          if(fabs(W->umax[i] - u_j) < eps) DQ = -W->q[i];
          if(F[pij] > 0) DP = D[pij];
        } else {
          if(fabs(W->umax[i] - u_j) >= eps) DQ = W->q[i];
          DP = -row_d_sum(L, W, i, +1, 0);
        }
        if(W->Pp[i] > 0.0) {
          const double val = DQ / W->Pp[i] - W->Qp[i] * DP / (W->Pp[i] * W->Pp[i]);
          if(fabs(val) < epsinv) sum += val * F[k];
        }
      } else {
        if(ck != j) {
          if(fabs(W->umin[ck] - u_j) < eps) DQ = -W->q[ck];
          if(F[pij] < 0) DP = D[pij];
        } else {
          if(fabs(W->umin[j] - u_j) >= eps) DQ = W->q[j];
          DP = -row_d_sum(L, W, j, -1, 0);
        }
        if(W->Pm[ck] < 0.0) {
          const double val = DQ / W->Pm[ck] - W->Qm[ck] * DP / (W->Pm[ck] * W->Pm[ck]);
          if(fabs(val) < epsinv) sum += val * F[k];
        }
      }
      continue;
    }
    if(F[k] < 0) {
      if(W->Rm[i] <= W->Rp[ck]) {
        if(i != j) {
          if(fabs(W->umin[i] - u_j) < eps) DQ = -W->q[i];
          if(F[pij] < 0) DP = D[pij];
        } else {
          if(fabs(W->umin[i] - u_j) >= eps) DQ = W->q[i];
          DP = -row_d_sum(L, W, i, -1, 0);
        }
        if(W->Pm[i] < 0.0) {
          const double val = DQ / W->Pm[i] - W->Qm[i] * DP / (W->Pm[i] * W->Pm[i]);
          if(fabs(val) < epsinv) sum += val * F[k];
        }
      } else {
        if(ck != j) {
          if(fabs(W->umax[ck] - u_j) < eps) DQ = -W->q[ck];
          if(F[pij] > 0) DP = D[pij];
        } else {
          if(fabs(W->umax[j] - u_j) >= eps) DQ = W->q[j];
          DP = -row_d_sum(L, W, j, +1, 0);
        }
        if(W->Pp[ck] > 0.0) {
          const double val = DQ / W->Pp[ck] - W->Qp[ck] * DP / (W->Pp[ck] * W->Pp[ck]);
          if(fabs(val) < epsinv) sum += val * F[k];
        }
      }
    }
  }
  return sum;
}

/* Newton scheme (steady_state_algorithm :2794-2981), after the limiter:
   rhs_i = f_i + sum_j (-E_ij u_j + alpha_ij F_ij); the matrix
   DF_ij = E_ij - tau0 alpha_ij d_ij - tau1 J_ij off the diagonal,
   DF_ii = E_ii + tau0 sum_{j != i} alpha_ij d_ij - tau1 J_ii, J the
   limiter's Jacobian times the fluxes; then rhs_i += sum_j DF_ij u_j and
   the system matrix is DF (Dirichlet rows identity / u). tau1 = 0 skips J
   (its product with 0 changes nothing). */
static void afc_newton(olevel* L, oafc* W, const double* u, double* rhs, double tau0, double tau1, int bjk) {
  const int n = L->ndof, na = L->nactive;
  const double *E = W->E, *D = W->D;
  for(int i = 0; i < na; ++i) {
    double r = W->f[i];
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) r += -(E[l]) * u[L->col[l]] + W->alpha[l] * W->F[l];
    rhs[i] = r;
  }
  for(int i = 0; i < na; ++i)
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      const int c = L->col[l];
      double J = 0;
      if(tau1 != 0 && orc_fault != 3) J = bjk ? jac_bjk(L, W, u[c], i, c, l) : jac_kuzmin(L, W, i, c, l);
      if(c != i) {
        W->df[l] = E[l] - tau0 * W->alpha[l] * D[l] - tau1 * J;
      } else {
        double sum_flux = 0.0;
        for(int jj = L->rp[i]; jj < L->rp[i + 1]; ++jj)
          if(L->col[jj] != i) sum_flux += W->alpha[jj] * D[jj];
        W->df[l] = E[l] + tau0 * sum_flux - tau1 * J;
      }
    }
  for(int i = 0; i < n; ++i) {
    if(i >= na) {
      for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) L->val[l] = L->col[l] == i ? 1.0 : 0.0;
      rhs[i] = u[i];
      continue;
    }
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
      rhs[i] += W->df[l] * u[L->col[l]];
      L->val[l] = W->df[l];
    }
  }
  extract_diag(L);
  if(L->lu) { free(L->lu); L->lu = NULL; free(L->piv); L->piv = NULL; }
}

/* Anderson acceleration (ConvectionDiffusion_AFC.C:600-686, anderson.f90):
   with the columns R_1..R_k of the last k updates and X_1..X_k of the last
   k fixed point outputs, alpha = argmin |R alpha| subject to sum alpha = 1
   (DGGLSE with B = (1 ... 1), d = 1, c = 0), and the new iterate
   sum_j alpha_j X_j. Solved as DGGLSE does: the Householder reflector H with
   (1 ... 1) H = (0 ... 0 r), then a Householder QR of R H; x_k = 1 / r,
   the first k - 1 unknowns from the triangular system, alpha = H x.
   Returns 0 when R H has a zero pivot (the iterate is left unchanged). */
static int anderson_alpha(int n, int k, const double* R, double* alpha) {
  /* reflector of the ones row: v = (1 ... 1) - s e_k with s = -sqrt(k) */
  double* v = malloc(sizeof(double) * k);
  const double r = -sqrt((double)k);
  for(int j = 0; j < k; ++j) v[j] = 1.0;
  v[k - 1] -= r;
  double vv = 0;
  for(int j = 0; j < k; ++j) vv += v[j] * v[j];
  /* M = R H (n x k, column major): M = R - (R v) (2 / vv) v^T */
  double* M = malloc(sizeof(double) * (size_t)n * k);
  for(int i = 0; i < n; ++i) {
    double rv = 0;
    for(int j = 0; j < k; ++j) rv += R[(size_t)j * n + i] * v[j];
    for(int j = 0; j < k; ++j) M[(size_t)j * n + i] = R[(size_t)j * n + i] - rv * (2.0 / vv) * v[j];
  }
  /* Householder QR of M in place: upper triangle = T */
  double* w = malloc(sizeof(double) * n);
  int ok = 1;
  for(int c = 0; c < k && c < n; ++c) {
    double* x = M + (size_t)c * n;
    double nrm = 0;
    for(int i = c; i < n; ++i) nrm += x[i] * x[i];
    nrm = sqrt(nrm);
    if(nrm == 0) { ok = c >= k - 1; continue; }
    const double a = x[c] > 0 ? -nrm : nrm;
    for(int i = c; i < n; ++i) w[i] = x[i];
    w[c] -= a;
    double ww = 0;
    for(int i = c; i < n; ++i) ww += w[i] * w[i];
    for(int cc = c; cc < k; ++cc) {
      double* y = M + (size_t)cc * n;
      double d = 0;
      for(int i = c; i < n; ++i) d += w[i] * y[i];
      const double t = 2.0 * d / ww;
      for(int i = c; i < n; ++i) y[i] -= t * w[i];
    }
  }
  double* xv = malloc(sizeof(double) * k);
  if(ok) {
    xv[k - 1] = 1.0 / r;
    /* T11 x1 = -T12 x_k */
    for(int i = k - 2; i >= 0; --i) {
      double s = -M[(size_t)(k - 1) * n + i] * xv[k - 1];
      for(int j = i + 1; j < k - 1; ++j) s -= M[(size_t)j * n + i] * xv[j];
      const double piv = M[(size_t)i * n + i];
      if(piv == 0) { ok = 0; break; }
      xv[i] = s / piv;
    }
  }
  if(ok) {
    double vx = 0;
    for(int j = 0; j < k; ++j) vx += v[j] * xv[j];
    for(int j = 0; j < k; ++j) alpha[j] = xv[j] - (2.0 / vv) * vx * v[j];
  }
  free(v); free(M); free(w); free(xv);
  return ok;
}

static double afc_residual(const olevel* L, const double* u, const double* rhs, double* tmp) {
  memcpy(tmp, rhs, sizeof(double) * L->ndof);
  apply_scaled_add(L, u, tmp, -1.0);
  return sqrt(ddot(L->ndof, tmp, tmp));
}

/* prm: epsilon, maxit, damping, damping_max, damping_min, damping_max_global,
   increase, decrease, max_increase, max_decrease, min_tol.
   out: iterations, final residual, rejected steps, initial residual */
int orc_afc_solve(orc_problem* P, const osolver* s, const double* prm, double* out) {
  if(P->nlev != 1) return 3;
  olevel* L = lvl(P, -1);
  const int n = L->ndof, na = L->nactive, nnz = L->nnz;
  P->afc = 1;
  assemble_level(P, L, 0, L->m.ncells);
  P->afc = 0;
  oafc W;
  const int bjk = prm[11] == 1.0;
  if(bjk && !((L->m.shape == TRI || L->m.shape == TET) && L->nl == NV[L->m.shape])) {
    P->afc = 0;
    return 4;
  }
  double* buf = calloc((size_t)7 * nnz + 11 * (size_t)n, sizeof(double));
  W.A = buf; W.D = W.A + nnz; W.E = W.D + nnz; W.F = W.E + nnz; W.alpha = W.F + nnz;
  double* S = W.alpha + nnz;
  W.df = S + nnz;
  W.f = W.df + nnz; W.Pp = W.f + n; W.Pm = W.Pp + n; W.Qp = W.Pm + n; W.Qm = W.Qp + n; W.Rp = W.Qm + n;
  W.Rm = W.Rp + n;
  W.gamma = W.Rm + n;
  W.umax = W.gamma + n; W.umin = W.umax + n; W.q = W.umin + n;
  double* old = malloc(sizeof(double) * n);
  double* cur_sol = malloc(sizeof(double) * n);
  double* cur_rhs = malloc(sizeof(double) * n);
  double* tmp = malloc(sizeof(double) * n);
  for(int l = 0; l < nnz; ++l) W.alpha[l] = 1.0;
  memcpy(W.A, L->val, sizeof(double) * nnz);
  memcpy(W.f, L->rhs, sizeof(double) * n);
  if(bjk) {
    /* modify_system_matrix (:35-60): a_di = 0 for a Dirichlet column d of
       an active row i with a_id < 0 */
    for(int i = 0; i < na; ++i)
      for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) {
        const int c = L->col[l];
        if(c >= na && W.A[l] < 0.0) {
          const int t = transposed_pos(L, c, i);
          if(t < L->rp[c + 1]) W.A[t] = 0.0;
        }
      }
    afc_gamma(P, L, W.gamma);
  }
  afc_diffusion(L, W.A, W.D);
  for(int l = 0; l < nnz; ++l) W.E[l] = W.A[l] + W.D[l];
  memcpy(S, W.E, sizeof(double) * nnz);
  for(int i = na; i < n; ++i)
    for(int l = L->rp[i]; l < L->rp[i + 1]; ++l) S[l] = L->col[l] == i ? 1.0 : 0.0;
  memcpy(L->val, S, sizeof(double) * nnz);
  extract_diag(L);
  if(L->lu) { free(L->lu); L->lu = NULL; free(L->piv); L->piv = NULL; }
  /* initial iterate: zero, Dirichlet values from the rhs */
  memset(L->sol, 0, sizeof(double) * na);
  memcpy(L->sol + na, W.f + na, sizeof(double) * (n - na));
  const double eps = prm[0] * sqrt((double)n);
  const int maxit = (int)prm[1];
  double damp = prm[2], damp_max = prm[3];
  const double damp_min = prm[4], max_global = prm[5], inc = prm[6], dec = prm[7], max_inc = prm[8],
               max_dec = prm[9], min_tol = prm[10];
  void (*limiter)(const olevel*, oafc*, const double*, double*) = bjk ? afc_bjk_rhs : afc_kuzmin_rhs;
  const int fpm = prm[12] == 1.0, newton = prm[12] == 2.0;
  const double tau0 = prm[13];
  /* Newton: the derivative's weight is afc_fixed_point_derivative_weight
     times a factor that starts at 0 and follows the residual
     (ConvectionDiffusion_AFC.C:427-460) */
  const double deriv_weight = prm[14], threshold = prm[15];
  double deriv_factor = 0.0;
  /* Anderson acceleration: the last kA fixed point outputs and updates */
  const int anderson = prm[16] == 1.0, kA = (int)prm[17], startA = (int)prm[18];
  double *histX = NULL, *histR = NULL, *Rmat = NULL, *alphaA = NULL;
  int n_hist = 0;
  if(anderson) {
    histX = malloc(sizeof(double) * (size_t)n * kA);
    histR = malloc(sizeof(double) * (size_t)n * kA);
    Rmat = malloc(sizeof(double) * (size_t)n * kA);
    alphaA = malloc(sizeof(double) * kA);
  }
#define limited_rhs(L_, W_, u_, r_)                                          \
  do {                                                                       \
    limiter(L_, W_, u_, r_);                                                 \
    if(fpm) afc_fpm(L_, W_, u_, r_, tau0);                                   \
    if(newton) afc_newton(L_, W_, u_, r_, tau0, deriv_weight * deriv_factor, bjk); \
  } while(0)
  /* iteration 0 */
  limited_rhs(L, &W, L->sol, L->rhs);
  double residual_old = afc_residual(L, L->sol, L->rhs, tmp), residual = residual_old;
  out[3] = residual_old;
  memcpy(old, L->sol, sizeof(double) * n);
  ostats st;
  int rc = orc_solve(P, s, &st);
  int k = 1, rejected = 0;
  for(;; ++k) {
    const int acc = anderson && k >= startA;
    if(acc) {
      /* anderson_acceleration_damping: append the output and the update,
         drop the oldest beyond kA, mix once kA are there */
      if(n_hist == kA) {
        memmove(histX, histX + n, sizeof(double) * (size_t)n * (kA - 1));
        memmove(histR, histR + n, sizeof(double) * (size_t)n * (kA - 1));
        --n_hist;
      }
      for(int i = 0; i < n; ++i) {
        histX[(size_t)n_hist * n + i] = L->sol[i];
        histR[(size_t)n_hist * n + i] = L->sol[i] - old[i];
      }
      ++n_hist;
      if(k > kA + startA && n_hist == kA) {
        memcpy(Rmat, histR, sizeof(double) * (size_t)n * kA);
        if(orc_fault != 4 && anderson_alpha(n, kA, Rmat, alphaA))
          for(int i = 0; i < n; ++i) {
            double sum = 0.0;
            for(int j = 0; j < kA; ++j) sum = sum + histX[(size_t)j * n + i] * alphaA[j];
            L->sol[i] = sum;
          }
      }
    }
    double omega = damp;
    int first_damp = 1;
    memcpy(cur_rhs, L->rhs, sizeof(double) * n);
    memcpy(cur_sol, L->sol, sizeof(double) * n);
    for(;;) {
      for(int i = 0; i < n; ++i) L->sol[i] = old[i] + omega * (L->sol[i] - old[i]);
      memcpy(L->sol + na, old + na, sizeof(double) * (n - na));
      limited_rhs(L, &W, L->sol, L->rhs);
      residual = afc_residual(L, L->sol, L->rhs, tmp);
      if(k == 1) break;
      if(residual < residual_old || omega <= min_tol * damp_min) {
        if(residual < residual_old && first_damp == 1) {
          damp_max = fmin(max_global, max_inc * damp_max);
          omega = fmin(damp_max, inc * omega);
        }
        damp = omega;
        break;
      }
      if(!acc) {
        ++rejected;
        memcpy(L->sol, cur_sol, sizeof(double) * n);
        memcpy(L->rhs, cur_rhs, sizeof(double) * n);
      }
      omega = fmax(damp_min, omega * dec);
      if(first_damp == 1) {
        damp_max = fmax(damp_min, max_dec * damp_max);
        first_damp = 0;
      }
      damp = omega;
      /* with Anderson acceleration a rejected step only lowers the damping */
      if(acc) break;
    }
    if(residual < eps && k > 1) break;
    if(k == maxit) break;
    memcpy(old, L->sol, sizeof(double) * n);
    rc = orc_solve(P, s, &st);
    if(rc) break;
    if(newton) {
      if(residual <= threshold) {
        double w = deriv_factor;
        if(w < 1e-3) w = 0.25;
        if(residual / residual_old > 0.99) {
          w *= 0.999;
          if(w < 0.1) w = 0.1;
        }
        if(residual / residual_old < 0.99) {
          w *= 1.001;
          if(w > 1.0) w = 1.0;
        }
        deriv_factor = w;
      }
      if(residual > 100 * threshold) deriv_factor = 0.0;
    }
    residual_old = residual;
  }
#undef limited_rhs
  free(histX); free(histR); free(Rmat); free(alphaA);
  out[0] = k;
  out[1] = residual;
  out[2] = rejected;
  free(buf); free(old); free(cur_sol); free(cur_rhs); free(tmp);
  return rc;
}

void orc_tcd_matrices(orc_problem* P, int level, double* a, double* m) {
  olevel* L = lvl(P, level);
  if(a) memcpy(a, L->aval, sizeof(double) * L->nnz);
  if(m) memcpy(m, L->mval, sizeof(double) * L->nnz);
}

void orc_vcycle(orc_problem* P, const osolver* s, const double* v, double* z) {
  if(P->nlev > 1 && !P->lev[0].lu) dense_factor(&P->lev[0]);
  mg_cycle(P, s, v, z);
}

/* L2, H1-semi, SD, DG, Linf */
void orc_errors(orc_problem* P, double* out) {
  olevel* L = lvl(P, -1);
  const omesh* m = &L->m;
  const int nl = L->nl, d = P->el.dim;
  const oquad* q = &P->qd;
  static double ref[MAXQ][10 * MAXNL];
  for(int k = 0; k < q->n; ++k) elem_eval(&P->el, q->x[k], ref[k]);
  double err[5] = {0, 0, 0, 0, 0};
  double o[7][MAXNL];
  for(int c = 0; c < m->ncells; ++c) {
    ogeo g;
    geo(m, c, &g);
    double U[MAXNL], Le[5] = {0, 0, 0, 0, 0};
    for(int l = 0; l < nl; ++l) U[l] = L->sol[L->dof[c * nl + l]];
    for(int k = 0; k < q->n; ++k) {
      orig_values(P, &g, ref[k], o);
      double v[4] = {0, 0, 0, 0};
      for(int t = 0; t < 4; ++t)
        for(int l = 0; l < nl; ++l) v[t] += U[l] * o[t][l];
      double x[3], ex[4], cf[6];
      geo_point(&g, d, q->x[k], x);
      const double w = q->w[k] * fabs(g.det);
      ex_exact(&P->ex, x[0], x[1], d == 3 ? x[2] : 0, ex);
      ex_coeffs(&P->ex, x[0], x[1], d == 3 ? x[2] : 0, cf);
      const double nu = cf[0], b1 = cf[1], b2 = cf[2], b3 = d == 3 ? cf[3] : 0, cr = cf[4];
      const double delta = sdfem_delta(d, g.hK, nu, b1, b2, b3);
      const double e0 = v[0] - ex[0], e1 = v[1] - ex[1], e2 = v[2] - ex[2], e3 = d == 2 ? 0. : v[3] - ex[3];
      double e4 = b1 * e1 + b2 * e2;
      if(d == 3) e4 += b3 * e3;
      Le[0] += w * e0 * e0;
      if(fabs(e0) > Le[4]) Le[4] = fabs(e0);
      Le[1] += w * e1 * e1;
      Le[1] += w * e2 * e2;
      Le[1] += w * e3 * e3;
      Le[2] += w * (nu * (e1 * e1 + e2 * e2 + e3 * e3) + cr * e0 * e0 + delta * e4 * e4);
      Le[3] += w * (nu * (e1 * e1 + e2 * e2 + e3 * e3) + (e4 + cr * e0) * e0);
    }
    for(int k = 0; k < 4; ++k) err[k] += Le[k];
    if(err[4] < Le[4]) err[4] = Le[4];
  }
  for(int k = 0; k < 4; ++k) out[k] = sqrt(err[k]);
  out[4] = err[4];
}

/* standalone kernels on caller-provided CSR data (reference semantics) */
void orc_spmv_csr(int n, int n_active, const int* rp, const int* col, const double* val, const double* x,
                  double* y, double a) {
  for(int i = 0; i < n_active; ++i) {
    double v = 0;
    for(int j = rp[i]; j < rp[i + 1]; ++j) v += val[j] * x[col[j]];
    y[i] += a * v;
  }
  for(int i = n_active; i < n; ++i) y[i] += a * x[i];
}

double orc_ddot(int n, const double* x, const double* y) { return ddot(n, x, y); }

/* basis values at a reference point (for closed-form checks) */
void orc_basis(int shape, int order, const double* p, double* out) {
  oelem e = make_elem(shape, order);
  elem_eval(&e, p, out);
}

int orc_quad(int shape, int degree, double* w, double* x) {
  oquad q = make_quad(shape, degree);
  for(int i = 0; i < q.n; ++i) {
    w[i] = q.w[i];
    for(int k = 0; k < 3; ++k) x[3 * i + k] = q.x[i][k];
  }
  return q.n;
}

/* timing samples for the CPU baseline: assemble cells [c0, c1) of the finest
   level into its (zeroed) matrix, and run `iters` FGMRES iterations */
void orc_assemble_cells(orc_problem* P, int c0, int c1) {
  olevel* L = lvl(P, -1);
  if(c1 > L->m.ncells) c1 = L->m.ncells;
  assemble_level(P, L, c0, c1);
}

/* load a matrix / rhs computed elsewhere (same sparsity) into a level, e.g.
   to time the solver phase on exactly the system another path assembled */
void orc_set_matrix(orc_problem* P, int level, const double* v) {
  olevel* L = lvl(P, level);
  memcpy(L->val, v, sizeof(double) * L->nnz);
  extract_diag(L);
  if(L->lu) { free(L->lu); L->lu = NULL; free(L->piv); L->piv = NULL; }
}
void orc_set_rhs(orc_problem* P, int level, const double* v) {
  olevel* L = lvl(P, level);
  memcpy(L->rhs, v, sizeof(double) * L->ndof);
  if(L == lvl(P, -1)) {
    memset(L->sol, 0, sizeof(double) * L->nactive);
    memcpy(L->sol + L->nactive, L->rhs + L->nactive, sizeof(double) * (L->ndof - L->nactive));
  }
}

/* TMatrix::sor_sweep on a level's matrix or on any CSR matrix (in place) */
void orc_sor_sweep(orc_problem* P, int level, const double* b, double* x, double omega, int flag) {
  sor_level(lvl(P, level), b, x, omega, flag);
}
void orc_sor_csr(int n, const int* rp, const int* col, const double* val, const double* b, double* x,
                 double omega, int flag) {
  sor_csr(n, rp, col, val, b, x, omega, flag);
}

/* residual history of the last FGMRES/CG solve (LoopInfo output) */
int orc_history(double* out, int max) {
  int n = g_nhist < max ? g_nhist : max;
  memcpy(out, g_hist, sizeof(double) * n);
  return n;
}
