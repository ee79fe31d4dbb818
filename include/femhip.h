/*
 * femhip -- C ABI of the MI355X assembly library for the scikit-fem cell
 * assembly path (BilinearForm.assemble / LinearForm.assemble on CellBasis).
 *
 * Conventions
 *   - Every pointer named in a femhip_*_t struct or passed as an argument is
 *     a device pointer owned by the caller unless the comment says "host".
 *   - Layouts are the reference's: node coordinates p are [dim][nnodes]
 *     (row-major, i.e. structure-of-arrays), connectivity t is
 *     [nverts][nelems], element_dofs is [nbfun][nelems] (row = local index).
 *   - All floating point work is FP64, all indices are int32 (the row/col
 *     keys of the symbolic phase are int64 internally).
 *   - Functions enqueue on the given stream and do not synchronize unless the
 *     comment says so.  Return value 0 = success, negative = FEMHIP_E*;
 *     femhip_last_error() then holds a message (thread-local).
 */
#ifndef FEMHIP_H
#define FEMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* femhip_stream_t; /* a hipStream_t */

enum {
  FEMHIP_OK = 0,
  FEMHIP_EINVAL = -1,       /* invalid argument (ValueError)            */
  FEMHIP_EUNSUPPORTED = -2, /* element/form/quadrature not supported     */
  FEMHIP_EOVERFLOW = -3,    /* int32 index range exceeded                */
  FEMHIP_EHIP = -4,         /* HIP runtime error                         */
  FEMHIP_EDEGENERATE = -5,  /* zero Jacobian determinant (isoparametric) */
  FEMHIP_ESIZE = -6         /* caller buffer or workspace too small      */
};

enum { FEMHIP_TRI = 0, FEMHIP_TET = 1, FEMHIP_HEX = 2 };

enum {
  FEMHIP_LAPLACE = 0,       /* dot(grad u, grad v)                      */
  FEMHIP_MASS = 1,          /* u * v                                    */
  FEMHIP_KAPPA_LAPLACE = 2, /* dot(w.kappa * grad u, grad v)            */
  FEMHIP_ELASTICITY = 3,    /* ddot(C(sym_grad u), sym_grad v)          */
  FEMHIP_UNIT_LOAD = 4,     /* v  (linear form)                         */
  FEMHIP_COEF_MASS = 5,     /* w.c * u * v  (Robin / weighted mass)     */
  FEMHIP_COEF_LOAD = 6      /* w.c * v  (Neumann / source data)         */
};

/* Mesh geometry and quadrature (CellBasis.mapping + quadrature). */
typedef struct femhip_geom {
  int32_t kind;         /* FEMHIP_TRI / TET (affine) or HEX (isoparametric) */
  int32_t dim;          /* 2 or 3                                          */
  int32_t nverts;       /* rows of t: 3, 4 or 8                            */
  int32_t nq;           /* quadrature points                               */
  int64_t nelems;
  int64_t nnodes;       /* columns of p                                    */
  const double* p;      /* [dim][nnodes]                                   */
  const int32_t* t;     /* [nverts][nelems]                                */
  const double* W;      /* [nq] quadrature weights                         */
  const double* gdphi;  /* HEX: [8][3][nq] Hex1 reference gradients at the
                           quadrature points (the mapping's element);
                           NULL for simplices                              */
} femhip_geom_t;

/* Finite element space (CellBasis.elem + Dofs). */
typedef struct femhip_space {
  int32_t nbs;          /* scalar basis functions per element              */
  int32_t vector;       /* 1 = ElementVector (dim components)              */
  int32_t nbfun;        /* nbs * (vector ? dim : 1)                        */
  int32_t reserved;
  int64_t ndofs;        /* N                                               */
  const int32_t* element_dofs; /* [nbfun][nelems]                          */
  const double* phi;    /* [nbs][nq]      reference values at X            */
  const double* dphi;   /* [nbs][dim][nq] reference gradients at X         */
} femhip_space_t;

/* Built-in form and its parameters (skfem.models). */
typedef struct femhip_form {
  int32_t kind;         /* FEMHIP_LAPLACE ...                              */
  int32_t kappa_mode;   /* 0: none, 1: kappa[e], 2: kappa[e*nq + q]        */
  const double* kappa;  /* the coefficient of KAPPA_LAPLACE, COEF_MASS,    */
                        /* COEF_LOAD                                       */
  double lam;           /* linear_elasticity Lambda                        */
  double two_mu;        /* 2. * Mu, as the reference evaluates it          */
} femhip_form_t;

const char* femhip_last_error(void);
int femhip_version(void);

/* ---------------------------------------------------------------------
 * Mapping.  Replaces MappingAffine._init_Ab/_init_invA/detDF/invDF
 * (skfem/mapping/mapping_affine.py:55-131,205-232) and
 * MappingIsoparametric._J/detDF/invDF
 * (skfem/mapping/mapping_isoparametric.py:112-226), evaluated in the
 * reference's operation order.  Output G is
 *   affine: [D*D+1][nelems]        (invA row-major, then |detA|)
 *   hex:    [10][nelems*nq]        (invDF row-major, then |detDF|)
 * femhip_geometry_size returns the number of doubles of G.
 * --------------------------------------------------------------------- */
int64_t femhip_geometry_size(const femhip_geom_t* g);
/* status: device int32, set to 1 when an isoparametric detDF is exactly 0
 * (the reference raises "Zero Jacobian determinant"); may be NULL. */
int femhip_geometry(const femhip_geom_t* g, double* G, int32_t* status,
                    femhip_stream_t s);

/* ---------------------------------------------------------------------
 * Local element matrices / vectors in the reference's COO data order.
 * Replaces BilinearForm._assemble + _kernel
 * (skfem/assembly/form/bilinear_form.py:58-128,150-151) and
 * LinearForm._assemble (skfem/assembly/form/linear_form.py:18-49), with
 * the CellBasis precompute (cell_basis.py:101-105, element_h1.py:10-24,
 * element_vector.py:36-48) fused in.  Values are bit-identical to the
 * reference's NumPy evaluation when phi/dphi/W are (no FMA contraction,
 * einsum order, NumPy pairwise quadrature sum).
 *   bilinear: contrib[(j*nbfun + i)*nelems + e], j trial, i test
 *   linear:   contrib[i*nelems + e]
 * --------------------------------------------------------------------- */
int femhip_local_bilinear(const femhip_geom_t* g, const femhip_space_t* sp,
                          const femhip_form_t* f, const double* G,
                          double* contrib, femhip_stream_t s);
/* Trial functions from u, test functions from v (BilinearForm._assemble
 * with vbasis != ubasis, bilinear_form.py:58-128): both spaces tabulated at
 * the same quadrature points of the same elements, both scalar or both
 * ElementVector; the mapping and weights are the trial basis's (g, G).
 *   contrib[(j*v->nbfun + i)*nelems + e], j trial, i test; rows are v's
 *   DOFs, columns u's (shape (v->ndofs, u->ndofs)). */
int femhip_local_bilinear_mixed(const femhip_geom_t* g, const femhip_space_t* u,
                                const femhip_space_t* v, const femhip_form_t* f,
                                const double* G, double* contrib, femhip_stream_t s);
int femhip_local_linear(const femhip_geom_t* g, const femhip_space_t* sp,
                        const femhip_form_t* f, const double* G,
                        double* contrib, femhip_stream_t s);

/* Boundary facets of a P1/P2 simplex mesh (FacetBasis, facet_basis.py:23-110):
 * facets[dim][nfacets] their vertices (Mesh.facets order), tind[nfacets]
 * the owning element, X[dim-1][nq] / W[nq] the facet quadrature rule. */
typedef struct femhip_facet_geom {
  int32_t kind;         /* FEMHIP_TRI or FEMHIP_TET                       */
  int32_t dim;          /* 2 or 3                                         */
  int32_t degree;       /* 1 (P1) or 2 (P2)                               */
  int32_t nq;
  int64_t nfacets;
  int64_t nnodes;
  int64_t nelems;
  const double* p;      /* [dim][nnodes]                                  */
  const int32_t* facets;
  const int32_t* t;     /* [dim+1][nelems]                                */
  const int32_t* tind;
  const double* X;
  const double* W;
} femhip_facet_geom_t;

/* Local entries of a boundary form on the facets, in the reference's
 * arithmetic (FacetBasis.__init__, facet_basis.py:93-110: MappingAffine.G
 * mapping_affine.py:234-245, invF :195-203, detDG :247-253; the owning
 * element's invA from its vertices, :55-131):
 * form FEMHIP_MASS / FEMHIP_COEF_MASS -> out[(j*nb + i)*nfacets + f]
 * (BilinearForm._assemble layout, bilinear_form.py:58-128),
 * FEMHIP_UNIT_LOAD / FEMHIP_COEF_LOAD -> out[i*nfacets + f]
 * (linear_form.py:18-49).  The COEF forms read coef[f] (coef_mode 1) or
 * coef[f*nq + q] (coef_mode 2).  nq is 1..7. */
int femhip_facet_local(const femhip_facet_geom_t* g, int32_t form, const double* coef,
                       int32_t coef_mode, double* out, femhip_stream_t s);

/* FacetBasis.interpolate (abstract_basis.py:271-322 on a FacetBasis) of a
 * DOF vector w: value[f*nq + q] and grad[(j*nfacets + f)*nq + q] (grad may be
 * NULL); dofs[i*nfacets + f] = element_dofs[i, tind[f]]. */
int femhip_facet_interpolate(const femhip_facet_geom_t* g, const int32_t* dofs, const double* w,
                             double* value, double* grad, femhip_stream_t s);

/* FacetBasis.global_coordinates / mesh_parameters (facet_basis.py:126-139):
 * x[(i*nfacets + f)*nq + q] = G(X) (mapping_affine.py:234-245) and
 * h[f*nq + q] = |detB| ** (1/(dim-1)) (h may be NULL). */
int femhip_facet_global_coordinates(const femhip_facet_geom_t* g, double* x, double* h,
                                    femhip_stream_t s);

/* Outward unit normals of the facets seen from tind_normals[f] (= f2t[0]),
 * whose local facet (RefTri/RefTet order) is floc[f]
 * (MappingAffine.normals, mapping_affine.py:255-281; FacetBasis.normals,
 * facet_basis.py:96-104): out[(j*nfacets + f)*nq + q]. */
int femhip_facet_normals(const femhip_facet_geom_t* g, const int32_t* tind_normals,
                         const int32_t* floc, double* out, femhip_stream_t s);

/* CellBasis.global_coordinates / mesh_parameters (cell_basis.py:129-141):
 * x[(i*nelems + e)*nq + q] = F(X_q) (MappingAffine.F, mapping_affine.py:183-192,
 * from X[dim][nq]; MappingIsoparametric.F, mapping_isoparametric.py:50-64,
 * from phi1[8][nq], the trilinear basis at the points) and
 * h[e*nq + q] = |detDF| ** (1/dim) from the mapping G of femhip_geometry. */
int femhip_global_coordinates(const femhip_geom_t* g, const double* X, const double* phi1,
                              double* x, femhip_stream_t s);
int femhip_mesh_parameters(const femhip_geom_t* g, const double* G, double* h,
                           femhip_stream_t s);

/* A scalar CSR matrix (n rows, nnz entries) expanded to the components of
 * ElementVector (element_vector.py:36-48): row dim*r + c holds scalar row r
 * with columns dim*s + c and the same values (data may be NULL for the
 * pattern alone).  The assembly of vector_laplace (models/poisson.py:12-14)
 * on ElementVector(e) is the expansion of laplace on e: the same values and
 * the same exact zeros.  out: indptr[dim*n + 1], indices/data[dim*nnz]. */
int femhip_csr_expand_components(const int32_t* indptr, const int32_t* indices,
                                 const double* data, int64_t n, int64_t nnz, int32_t dim,
                                 int32_t* out_indptr, int32_t* out_indices, double* out_data,
                                 femhip_stream_t s);

/* Basis.interpolate (abstract_basis.py:271-322) of a DOF vector w[ndofs]
 * for a scalar element: value[nelems][nq] and, when grad != NULL,
 * grad[dim][nelems][nq] (the global gradients of ElementH1.gbasis,
 * element_h1.py:10-24, G from femhip_geometry), bit-identical to the
 * reference's values. */
int femhip_interpolate(const femhip_geom_t* g, const femhip_space_t* sp, const double* G,
                       const double* w, double* value, double* grad, femhip_stream_t s);

/* ---------------------------------------------------------------------
 * Symbolic phase.  Replaces coo_matrix -> eliminate_zeros -> tocsr
 * (skfem/assembly/form/coo_data.py:28-36) for the pattern, and
 * COOData.toarray (coo_data.py:102-108) for the load-vector indexing.
 *
 * femhip_pattern_build: contributions k = (j*nb_col... ) as above; a
 * contribution is dropped when contrib[k] == 0.0 (eliminate_zeros runs
 * before duplicates are summed).  contrib == NULL keeps all of them.
 * Outputs: indptr[nrows+1], indices[cap ncontrib], perm[cap ncontrib]
 * (contribution ids sorted by (row, col), stable), seg[cap ncontrib+1]
 * (perm range of every CSR slot), *nnz_out (host; this call synchronizes
 * the stream).  Workspace size from femhip_pattern_workspace.
 *
 * femhip_rows_plan_build: the same for a vector (rows only, no pruning):
 * perm[ncontrib], seg[nrows+1].
 * --------------------------------------------------------------------- */
int femhip_pattern_workspace(int64_t ncontrib, size_t* bytes);
int femhip_pattern_build(const int32_t* row_dofs, const int32_t* col_dofs,
                         int32_t nb_row, int32_t nb_col, int64_t nelems,
                         int64_t nrows, int64_t ncols, const double* contrib,
                         int32_t* indptr, int32_t* indices, int32_t* perm,
                         int32_t* seg, int64_t* nnz_out, void* workspace,
                         size_t ws_bytes, femhip_stream_t s);
int femhip_rows_plan_build(const int32_t* row_dofs, int32_t nb_row,
                           int64_t nelems, int64_t nrows, int32_t* perm,
                           int32_t* seg, void* workspace, size_t ws_bytes,
                           femhip_stream_t s);

/* Exact-zero mask of contrib[n] (the contributions eliminate_zeros drops,
 * coo_data.py:34): bit (c & 63) of mask[c >> 6] is set iff contrib[c] ==
 * 0.0; mask holds ceil(n / 64) words (device).  A plan built from one
 * contrib array serves another one exactly when their masks are equal. */
int femhip_zero_mask(const double* contrib, int64_t n, uint64_t* mask,
                     femhip_stream_t s);
/* *count (device int64) <- number of contributions of contrib[n] whose
 * zero bit differs from mask (0: the same CSR pattern). */
int femhip_zero_mask_mismatch(const double* contrib, int64_t n,
                              const uint64_t* mask, int64_t* count,
                              femhip_stream_t s);

/* Numeric phase of the generic path: out[k] = sum of contrib[perm[j]] for
 * j in [seg[k], seg[k+1]), summed in perm order (deterministic). */
int femhip_segment_sum(const double* contrib, const int32_t* perm,
                       const int32_t* seg, int64_t nseg, double* out,
                       femhip_stream_t s);

/* ---------------------------------------------------------------------
 * Fused P1 path (triangles / tetrahedra; laplace, mass, kappa_laplace).
 * One pass replaces BilinearForm._assemble + COOData._assemble_scipy_csr
 * (bilinear_form.py:58-148, coo_data.py:28-36) once the CSR pattern is
 * known: every workgroup owns a set of rows (a spatial node partition),
 * recomputes the local matrices of the elements touching them in
 * registers, sums the contributions in LDS as fixed-point integers (one
 * exponent per launch, 2^46 <= max|contribution| 2^s < 2^50; error <=
 * n * 7.2e-15 * max|contribution| for a slot with n contributions;
 * deterministic) and writes each CSR row once.
 *
 * femhip_fused_plan_build runs on the host (host pointers: t, p, indptr,
 * indices of the pattern) and keeps the plan in library-owned host memory
 * until femhip_fused_plan_free; femhip_fused_plan_export copies its arrays
 * into caller buffers sized from the info struct:
 *   parts[8*nparts] nodemap[nlocal] rows[2*nrows] recs[rec_words*nrecs]
 *   rec_elem[nrecs].
 * order: 0 = records of a partition in greedy colour classes, 1 = in
 * element order.
 * --------------------------------------------------------------------- */
typedef struct femhip_fused_plan_info {
  int32_t nparts, max_recs, nrows, max_slots;
  int32_t max_nodes, max_rows, rec_words, idbits;
  int32_t offbits, block;
  int32_t pruned;       /* 1: the pattern lacks a pair of some element      */
  int32_t nrowrecs;     /* rows plans: (visit, owned DOF) records           */
  int64_t nrecs, nlocal;
  double dup;           /* element visits / elements (halo recomputation)  */
} femhip_fused_plan_info_t;

int femhip_fused_plan_build(int32_t nverts, int32_t dim, int64_t nelems, int64_t nnodes,
                            const int32_t* t, const double* p, const int32_t* indptr,
                            const int32_t* indices, int32_t slot_cap, int32_t node_cap,
                            int32_t block, int32_t order, void** handle,
                            femhip_fused_plan_info_t* info);
int femhip_fused_plan_export(void* handle, int32_t* parts, int32_t* nodemap, int32_t* rows,
                             uint32_t* recs, int32_t* rec_elem);
void femhip_fused_plan_free(void* handle);

/* Rows plan of a P2 simplex mesh (nverts 6 / 10 DOFs per element, `p` the
 * DOF locations used only for the partition): the same partitions, node
 * maps and row tables as a fused plan of the DOF graph, records holding the
 * element's local DOF ids (16 bits each) followed by the element id (one
 * word; info.rec_words counts it) and one row record per (visit,
 * owned DOF) of six words -- (visit << 4) | local DOF, then 7-bit column
 * offsets four per word: of the element's DOFs in the DOF's row, and of the
 * DOF in the rows of the element's other owned DOFs (127: not in the
 * pattern), the DOF's local row in the top nibbles of words 1-4 --
 * exported by femhip_rows_fused_plan_export (info.nrowrecs records;
 * parts[8k + 5..7] = first row record, row records, first visit).  Freed with
 * femhip_fused_plan_free; the other arrays by femhip_fused_plan_export. */
int femhip_rows_fused_plan_build(int32_t nverts, int32_t dim, int64_t nelems, int64_t ndofs,
                                 const int32_t* t, const double* p, const int32_t* indptr,
                                 const int32_t* indices, int32_t slot_cap, int32_t node_cap,
                                 int32_t block, void** handle,
                                 femhip_fused_plan_info_t* info);
int femhip_rows_fused_plan_export(void* handle, int32_t* rowrecs);


/* The exported plan in device memory. */
typedef struct femhip_fused_dev {
  const int32_t* parts;
  const int32_t* nodemap;
  const int32_t* rows;
  const uint32_t* recs;
  const int32_t* rec_elem;
  femhip_fused_plan_info_t info;
} femhip_fused_dev_t;

/* Accumulation and validation state of one form on one plan, in device
 * memory (256 bytes, 128-byte aligned as every hipMalloc / torch allocation
 * is; zeroed before first use, then owned by the kernels; calls that share
 * a state must not overlap).  Two cache lines: the first holds what every
 * workgroup of a pass reads, the second what they update atomically -- a
 * line that takes device-wide atomics is slow to read with plain loads. */
enum { FEMHIP_FUSED_BAD = 1, FEMHIP_FUSED_OVERFLOW = 2 };
typedef struct femhip_fused_state {
  /* read by every workgroup of a first pass */
  int32_t s, sv;         /* fixed-point exponents for the next first pass     */
  int32_t valid;         /* s, sv have been measured                          */
  int32_t s_first;       /* exponents the last first pass used (INT32_MIN:    */
  int32_t sv_first;      /*   none)                                           */
  int32_t reserved0;
  uint64_t epoch;        /* tag of the next first pass's maxima               */
  uint32_t reserved1[24];
  /* updated atomically */
  uint64_t maxk, maxv;   /* (epoch << 11) | exponent field of the pass's largest
                            |contribution| / load entry (atomicMax)           */
  int32_t flags;         /* FEMHIP_FUSED_BAD | FEMHIP_FUSED_OVERFLOW           */
  int32_t checked;       /* workgroups of the recheck that have finished      */
  int64_t nsusp;         /* suspects queued by the last first pass            */
  int64_t mismatch, zeros; /* recheck counters                                */
  uint32_t reserved3[20];
} femhip_fused_state_t;

/* Validation of a pass against the pattern the plan was built for (the
 * changed-input check): entries within the rounding bound of zero are
 * queued and recomputed in the reference's arithmetic; the pass is valid
 * iff exactly the plan's zero contributions (its femhip_zero_mask, nzeros
 * set bits) are zero.  verdict (device int32[4]) <- {valid, mismatches,
 * zeros found, flags}.  queue: device uint64[capacity] scratch. */
typedef struct femhip_fused_check {
  const int32_t* t;        /* connectivity [nverts][nelems] (reference layout) */
  const double* dphi;      /* reference gradients [nverts][dim][Q]             */
  const double* W;         /* quadrature weights [Q]                           */
  int64_t nelems;
  int32_t Q, reserved;
  const uint64_t* zero_mask;
  int64_t nzeros;
  uint64_t* queue;
  int64_t capacity;
  int32_t* verdict;
} femhip_fused_check_t;

/* LDS bytes per workgroup that femhip_fused_assemble asks for. */
size_t femhip_fused_lds_bytes(const femhip_fused_plan_info_t* info, int32_t dim,
                              int32_t with_load);

/* data[nnz] <- assembled CSR values of `form` (FEMHIP_LAPLACE, FEMHIP_MASS,
 * FEMHIP_KAPPA_LAPLACE with kappa[nelems] per element) on the pattern the
 * plan was built for; if load != NULL also load[nnodes] <- the unit-load
 * vector (LinearForm unit_load, linear_form.py:18-49) in the same pass.
 * p is [dim][nnodes] on the device.  Two launches (the second repeats the
 * pass only when the first measured another fixed-point exponent), then,
 * with check != NULL, the recheck of the suspects (verdict on the device).
 * groups = a persistent grid of that many workgroups walking the
 * partitions with a fixed stride (0: one workgroup per partition).  A plan may concatenate several meshes
 * (femhip.batch): then p, data and load are the concatenated arrays. */
int femhip_fused_assemble(const femhip_fused_dev_t* plan, femhip_fused_state_t* state,
                          int32_t form, int32_t dim, const double* p, int64_t nnodes,
                          const double* kappa, double* data, double* load,
                          const femhip_fused_check_t* check, int32_t groups,
                          femhip_stream_t s);

/* Laplace and mass (and, with load != NULL, the unit load) of a P1 basis
 * in one pass -- BilinearForm.assemble of laplace and of mass and
 * LinearForm.assemble of unit_load (bilinear_form.py:130-148,
 * linear_form.py:18-49) over one read of the records and coordinates.  The
 * plan is built on the mass pattern (info.pruned == 0); lrows[nrows][2]
 * gives, per plan row in the plan's row order, the row's start in the
 * Laplace pattern and the mask of its slots the Laplace pattern lacks (bit
 * k: slot k of the mass row), so laplace[] is written in the Laplace
 * pattern and mass[] in the plan's.  check validates the Laplace zeros as
 * for femhip_fused_assemble (the Laplace plan's zero mask). */
int femhip_fused_laplace_mass(const femhip_fused_dev_t* plan, const int32_t* lrows,
                              femhip_fused_state_t* state, int32_t dim, const double* p,
                              int64_t nnodes, double* laplace, double* mass, double* load,
                              const femhip_fused_check_t* check, int32_t groups,
                              femhip_stream_t s);

/* femhip_fused_assemble without any allocation, caller's or library's.  It
 * has no counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The accumulation states (one
 * per form) and the suspect queue live in a static device array of the
 * library (3 x 256 B + 65,536 suspect entries = 525,056 bytes per device,
 * held until the process ends), so the caller passes only the plan, the
 * inputs and the outputs (check: the tables, the zero mask and the
 * verdict; its queue and capacity fields are ignored).  Returns
 * FEMHIP_EINVAL for a check whose pattern has more exact zeros than the
 * static queue holds (femhip_fused_static_queue()).  Every call on a device
 * shares that device's array and keeps the last call's contents; calls
 * that overlap on one device race. */
int64_t femhip_fused_static_queue(void);
int femhip_fused_assemble_static(const femhip_fused_dev_t* plan, int32_t form, int32_t dim,
                                 const double* p, int64_t nnodes, const double* kappa,
                                 double* data, double* load, const femhip_fused_check_t* check,
                                 int32_t groups, femhip_stream_t s);

/* LDS bytes per workgroup that femhip_rows_fused_assemble asks for. */
size_t femhip_rows_fused_lds_bytes(const femhip_fused_plan_info_t* info);

/* Element metrics for femhip_rows_fused_assemble: metric[8 e + k] for the
 * affine simplex geometry g and its femhip_geometry G -- |det| A^-1 A^-T
 * (entries (0,0), (1,1), (2,2), (0,1), (0,2), (1,2); 2-D: (0,0), (1,1),
 * (0,1)) in k < D(D+1)/2, the Laplace suspect scale 2^-44 D max|A^-1|^2
 * |det| in k = 6 and |det| in k = 7.  metric: 8 * nelems doubles, 16-byte
 * aligned.  Part of MappingAffine (mapping_affine.py:55-131,205-232) as the
 * fused P2 kernel uses it. */
int femhip_rows_metric(const femhip_geom_t* g, const double* G, double* metric,
                       femhip_stream_t s);

/* data[nnz] <- CSR values of `form` (FEMHIP_LAPLACE, FEMHIP_MASS, or
 * FEMHIP_KAPPA_LAPLACE with kappa[nelems] per element; kappa NULL for the
 * others) for a P2 simplex basis (g, sp as for femhip_local_bilinear;
 * metric its femhip_rows_metric) on the pattern the rows plan was built for
 * (BilinearForm._assemble + COOData._assemble_scipy_csr,
 * bilinear_form.py:58-148, coo_data.py:28-36): every entry from the
 * element metric |det| A^-1 A^-T and tables of the form, summed in LDS rows,
 * no local-matrix array.  With check != NULL the call is validated against
 * the pattern as femhip_fused_assemble's is: entries within the rounding
 * bound of zero are queued (check->queue, check->capacity), recomputed in
 * the reference's arithmetic and compared with check->zero_mask /
 * check->nzeros (femhip_zero_mask of the contributions the pattern was
 * built from); check->verdict (device int32[4]) <- {valid, mismatches,
 * zeros, flags}; check's t/dphi/W fields are not used (g and sp give
 * them).  Two launches, as femhip_fused_assemble, then the recheck and the
 * verdict.  rowrecs: the exported row records. */
int femhip_rows_fused_assemble(const femhip_fused_dev_t* plan, const uint32_t* rowrecs,
                               femhip_fused_state_t* state, int32_t form,
                               const femhip_geom_t* g, const femhip_space_t* sp,
                               const double* metric, const double* kappa, double* data,
                               const femhip_fused_check_t* check, femhip_stream_t s);

/* out[S][nseg] <- kappa-weighted segmented sums  sum_k kappa[s][elem[k]] *
 * K1p[k]  over the plan's slots (seg): many per-element coefficient samples
 * (kappa[S][E]) assembled on one shared pattern from the kappa = 1 local
 * entries in plan order, K1p[k] = K1[perm[k]], elem[k] = perm[k] % E
 * (perm, seg: a femhip_pattern_build plan).  Batched counterpart of
 * BilinearForm.assemble (bilinear_form.py:130-148) called once per sample. */
int femhip_batched_kappa_values(const double* K1p, const int32_t* elem, const int32_t* seg,
                                int64_t nseg, int64_t nelems, const double* kappa,
                                int64_t nsamples, double* out, femhip_stream_t s);

/* Dirichlet rows on a device CSR (utils.enforce, skfem/utils.py:327-400):
 * rows D (distinct) are zeroed, their diagonal set to `diag`, b[D] = x[D]
 * (b, x may be NULL; x NULL means 0).  `missing` (device int32, may be
 * NULL) is set to 1 when a row of D has no diagonal slot in the pattern --
 * the reference inserts one (setdiag); the caller then falls back to the
 * host function, which does. */
int femhip_enforce_rows(const int32_t* indptr, const int32_t* indices, double* data,
                        const int32_t* D, int64_t nD, double diag, double* b, const double* x,
                        int32_t* missing, femhip_stream_t s);

/* ---------------------------------------------------------------------
 * Adjoint of the P1 (kappa-)Laplace + unit-load assembly.  The reference
 * has no adjoint (its autodiff, skfem/autodiff, differentiates with
 * respect to the solution through JAX); this replaces finite differences
 * of BilinearForm.assemble / LinearForm.assemble (bilinear_form.py:58-148,
 * linear_form.py:18-49) with respect to per-element kappa and the node
 * coordinates Mesh.p.
 *
 * femhip_fused_adjoint: on a fused plan of the connectivity pattern (no
 * pruned pair: info.pruned == 0; indptr its row starts, diag_slot[nnodes]
 * the CSR slot of each row's diagonal), from grad_data[nnz] and
 * grad_load[nnodes] (either may be NULL): dkappa[nelems] (may be NULL) and
 * dX[dim][nnodes], the node sums formed in the partitions' LDS (fixed
 * point, one exponent per launch, deterministic).  kappa NULL means
 * kappa = 1.  state: a femhip_fused_state_t of its own (not the forward's).
 * Two launches, as femhip_fused_assemble.
 * --------------------------------------------------------------------- */
size_t femhip_fused_adjoint_lds_bytes(const femhip_fused_plan_info_t* info, int32_t dim);
int femhip_fused_adjoint(const femhip_fused_dev_t* plan, femhip_fused_state_t* state,
                         int32_t dim, const double* p, int64_t nnodes, const int32_t* indptr,
                         const int32_t* diag_slot, const double* kappa,
                         const double* grad_data, const double* grad_load, double* dkappa,
                         double* dX, femhip_stream_t s);

/* ---------------------------------------------------------------------
 * The assembled system on the device, for the solve.
 *
 * Row/column selection (utils.condense, utils.py:540-603: A[I][:, I] and
 * b[I] - A[I][:, D] @ x[D]): rows[nsel] are the kept rows, colmap[ncols]
 * the new index of every kept column (increasing along the kept set) or -1.
 * femhip_csr_select_pattern writes out_indptr[nsel+1] and returns the nnz
 * (synchronizes); femhip_csr_select_fill writes out_indices/out_data and,
 * when b is given, b_out[r] = b[rows[r]] - sum_{j dropped} A[rows[r], j] x[j]
 * (x may be NULL: zero).
 * --------------------------------------------------------------------- */
int femhip_csr_select_workspace(int64_t nsel, size_t* bytes);
int femhip_csr_select_pattern(const int32_t* indptr, const int32_t* indices, const int32_t* rows,
                              int64_t nsel, const int32_t* colmap, int32_t* out_indptr,
                              int64_t* nnz_out, void* workspace, size_t ws_bytes,
                              femhip_stream_t s);
int femhip_csr_select_fill(const int32_t* indptr, const int32_t* indices, const double* data,
                           const int32_t* rows, int64_t nsel, const int32_t* colmap,
                           const int32_t* out_indptr, int32_t* out_indices, double* out_data,
                           const double* b, const double* x, double* b_out, femhip_stream_t s);

/* y[n] <- A x for a square CSR matrix (16 lanes per row, fixed summation
 * order); the product scipy's cg applies (utils.py:121-169). */
int femhip_spmv(const int32_t* indptr, const int32_t* indices, const double* data, int64_t n,
                const double* x, double* y, femhip_stream_t s);

/* Jacobi-preconditioned conjugate gradients on the device: the iteration of
 * utils.solver_iter_pcg (scipy.sparse.linalg.cg with M = build_pc_diag(A),
 * utils.py:56-58,121-169).  x holds the initial guess and receives the
 * iterate; stops when ||r|| <= max(rtol ||b||, atol) (checked every
 * check_every iterations, one host synchronization each) or after maxiter
 * iterations; *iters and *resid report them.  Needs a nonzero diagonal. */
size_t femhip_pcg_workspace_size(int64_t n);
int femhip_pcg(const int32_t* indptr, const int32_t* indices, const double* data, int64_t n,
               const double* b, double* x, double rtol, double atol, int64_t maxiter,
               int32_t check_every, void* workspace, size_t ws_bytes, int64_t* iters,
               double* resid, femhip_stream_t s);

/* ---------------------------------------------------------------------
 * Vector Q2 hexahedron linear elasticity on FP64 matrix cores.  Numeric
 * counterpart of BilinearForm._assemble for linear_elasticity(lam, mu) on
 * ElementVector(ElementHex2) with 27 Gauss points
 * (bilinear_form.py:58-128, models/elasticity.py:35-53): out[e][i*81 + j]
 * (element-major, row-major; i test, j trial) from the mapping G of
 * femhip_geometry.  groups: persistent workgroups (0 = automatic).
 * femhip_segment_sum_em: the segmented sums of femhip_segment_sum for such
 * element-major, row-major local matrices of nb x nb entries; contribution
 * c = (j*nb + i)*nelems + e of the plan names entry (i, j) of element e.
 * --------------------------------------------------------------------- */
int femhip_hex2_elasticity_mfma(const femhip_geom_t* g, const femhip_space_t* sp, double lam,
                                double mu, const double* G, double* out, int32_t groups,
                                femhip_stream_t s);
int femhip_segment_sum_em(const double* loc, const int32_t* perm, const int32_t* seg,
                          int64_t nseg, int64_t nelems, int64_t nb, double* out,
                          femhip_stream_t s);
/* The same sums row by row (COOData._assemble_scipy_csr, coo_data.py:28-36,
 * on element-major row-major local matrices): femhip_row_sum_em_plan writes
 * pos[(e*nb + i)*nb + j] = offset of entry (i, j) of element e in its CSR
 * row (the caller fills pos with -1 first: pruned contributions keep it);
 * femhip_row_sum_em sums, for every row r, the local rows
 * rowe[rowe_ptr[r] .. rowe_ptr[r+1]) (lr = e*nb + i with element_dofs[i][e]
 * == r, increasing e) into out[indptr[r] ..] in that order (rows of at most
 * 384 slots: max_row_len). */
int femhip_row_sum_em_plan(const int32_t* perm, const int32_t* seg, const int32_t* indptr,
                           int64_t nrows, int64_t nelems, int64_t nb, int16_t* pos,
                           femhip_stream_t s);
int femhip_row_sum_em(const double* loc, const int16_t* pos, const int32_t* rowe_ptr,
                      const int32_t* rowe, const int32_t* indptr, int64_t nrows, int64_t nb,
                      int32_t max_row_len, double* out, femhip_stream_t s);

#ifdef __cplusplus
}
#endif

#endif /* FEMHIP_H */
