/*
 * fem355 -- C ABI of the MI355X (gfx950) hex8 FEM hot path.
 *
 * Every entry point takes DEVICE pointers (e.g. torch.Tensor.data_ptr()),
 * sizes as int64_t and a hipStream_t passed as void*.  Calls are enqueued on
 * that stream and return immediately unless stated otherwise.  Every entry
 * point returns an int status: FEM_OK (0) or a negative FEM_E_* code, with a
 * message available from fem_last_error() (thread-local).  No C++ exception
 * crosses this boundary.  The library never frees caller memory and never
 * allocates device memory inside an entry point: scratch space is passed in by
 * the caller, sized by the matching *_work_bytes() query.
 *
 * Index conventions (the reference's, READ_physics.py:42-61):
 *   node index i in [0, N), element e in [0, E);
 *   elem[e*8 + j]  = node index of local node j (ANSYS I..P order);
 *   sm_node (N+1)  = SmNode, block-row pointer of the node-pair pattern;
 *   re_node (nnzb) = ReNode, sorted column node indices per row;
 *   co_node (N)    = CoNode, co_elem (N) = CoElem;
 *   ne_ptr (N+1), ne_list: node -> element incidences, entry = e*8 + j,
 *                  sorted by (e, j) within each node (all (e, j) pairs,
 *                  including repeated nodes of a degenerate element).
 * Vectors are node-interleaved: v[3*i + p].
 * BSR values are (nnzb, 3, 3) row-major: vals[9*s + 3*p + q].
 */
#ifndef FEM355_H
#define FEM355_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEM_OK 0
#define FEM_E_INVALID (-1)   /* bad argument: null pointer, bad size, bad enum */
#define FEM_E_HIP (-2)       /* HIP runtime error (message has the HIP string) */
#define FEM_E_LIMIT (-3)     /* a mesh exceeds a documented kernel limit       */
#define FEM_E_NUMERIC (-4)   /* singular pivot / non-finite value detected     */

#define FEM_DTYPE_F64 0
#define FEM_DTYPE_F32 1
#define FEM_DTYPE_F32F64 2   /* FP32 matrix values, FP64 vectors (solver kernels) */

#define FEM_LAYOUT_BSR 0     /* vals[9*s + 3*p + q]                              */
#define FEM_LAYOUT_CSR 1     /* the reference's CSR: 9*Sm[i] + 3*p*Co[i] + 3*m + q
                                (FEM_model_csr.py:162)                          */

/* Thread-local message describing the last failure. */
const char* fem_last_error(void);
/* ABI version, bumped on any signature change. */
int fem_abi_version(void);
/* Number of visible HIP devices (0 when none). */
int fem_device_count(int* count);

/* ------------------------------------------------------------------------
 * Mesh/pattern descriptor: all pointers are device pointers.
 * Row groups (grp_*) partition the node rows into spatially compact sets
 * whose incident elements are staged together in LDS by the row-owner
 * kernels; grp_enodes holds the 8 node indices of every element of every
 * group (elem[grp_elems[t]], so the kernel reaches coordinates in one
 * indirection).  tasks are the one-time symbolic contribution lists of the
 * assembly kernel: per group, each task names up to two element
 * contributions (local element, local row node j, local column node k) of
 * one output block (row-local index << 8 | slot in the row), the number T of
 * adjacent lanes sharing that block, or a second output block for "dual"
 * tasks.  Task word 0: out_a | out_b << 16 (out_b = 0xFFFF: none); word 1:
 * c1 | c2 << 14 | log2(T) << 28 | dual << 31 with c = elem << 6 | j << 3 | k
 * (0x3FFF: none).
 * ---------------------------------------------------------------------- */
typedef struct fem_mesh {
  int64_t num_node;
  int64_t num_elem;
  int64_t nnzb;
  const double* coords;        /* N x 3 */
  const int32_t* elem;         /* E x 8, 16-byte aligned */
  const int32_t* sm_node;      /* N + 1 */
  const int32_t* re_node;      /* nnzb  */
  const int32_t* co_node;      /* N     */
  const int32_t* co_elem;      /* N     */
  const int32_t* ne_ptr;       /* N + 1 */
  const int32_t* ne_list;      /* ne_ptr[N] */
  int64_t num_groups;
  int32_t max_group_elems;     /* max over groups of the element count */
  int32_t max_group_rows;      /* max over groups of the row count     */
  int32_t max_group_slots;     /* max over groups of sum CoNode of its rows; never
                                  less than the true value (fem_hex8_assemble sizes its
                                  LDS-staged copy-out by it, so an understated value is
                                  undefined behaviour); INT32_MAX when unknown, which
                                  turns the staging off */
  int32_t max_group_tasks;     /* max over groups of the task count */
  const int32_t* grp_row_ptr;  /* G + 1 */
  const int32_t* grp_rows;     /* N, rows in group order */
  const int32_t* grp_elem_ptr; /* G + 1 */
  const int32_t* grp_elems;    /* sorted element ids per group */
  const int32_t* grp_enodes;   /* 8 per entry of grp_elems, 16-byte aligned */
  const int32_t* grp_task_ptr; /* G + 1 */
  const int32_t* tasks;        /* 2 int32 per task, 8-byte aligned */
  const int32_t* grp_row_rec;  /* row records of the row-owner load-vector and stress
                                  kernels, 12 int32 per position p of grp_rows, 16-byte
                                  aligned: node i = grp_rows[p], CoElem[i], its incidence
                                  count c, an offset o into grp_inc, then its first 8
                                  incidences as (index of the element within the group's
                                  element list) << 3 | local node, ascending element;
                                  incidences 8 .. c-1 are grp_inc[o .. o + c - 8) */
  const int32_t* grp_inc;      /* overflow incidences of rows with more than 8 */
} fem_mesh;

/* ---- a1: symbolic pattern (replaces READ_physics.py:42-61 pre_process) --- */

/* Scratch bytes needed by fem_pattern_count. */
int64_t fem_pattern_work_bytes(int64_t num_elem, int64_t num_node);

/* Builds CoElem, the node->element incidence list (sorted) and, per node, the
 * neighbour count CoNode and its exclusive scan SmNode.  Synchronises the
 * stream (the caller needs nnzb = sm_node[N] to size re_node).
 * ne_list must hold 8*E entries.  FEM_E_INVALID if an element entry is
 * outside [0, N) (checked before any scatter); FEM_E_LIMIT if a node has
 * more than 128 distinct neighbour nodes (the per-thread sort buffer). */
int fem_pattern_count(const int32_t* elem, int64_t num_elem, int64_t num_node,
                      int32_t* co_elem, int32_t* ne_ptr, int32_t* ne_list,
                      int32_t* co_node, int32_t* sm_node,
                      void* work, int64_t work_bytes, void* stream);

/* Writes ReNode: sorted unique neighbour node indices per row. */
int fem_pattern_fill(const int32_t* elem, int64_t num_node, const int32_t* ne_ptr,
                     const int32_t* ne_list, const int32_t* sm_node, int32_t* re_node,
                     void* stream);

/* ---- a2/a3: CSR index arrays of the reference (FEM_model_csr.py:163,
 *      168-173, 191, 196-201) ---------------------------------------------- */

/* RowK (3N+1) and ColK (9*nnzb) of InstallIndexK / InstallMatrixK. */
int fem_csr_index_K(const fem_mesh* mesh, int32_t* row_k, int32_t* col_k, void* stream);
/* RowM (3N+1) and ColM (3*nnzb) of InstallIndexM / InstallMatrixM. */
int fem_csr_index_M(const fem_mesh* mesh, int32_t* row_m, int32_t* col_m, void* stream);
/* ValM (3*nnzb) of the reference's CSR mass from the per-block scalars. */
int fem_csr_values_M(const fem_mesh* mesh, const double* m_blocks, double* val_m, void* stream);

/* ---- a4-a7: hex8 stiffness / mass assembly (replaces InstallMatrixK
 *      FEM_model_csr.py:150-165, InstallMatrixKL FEM_model_bsr.py:163-175,
 *      InstallMatrixM FEM_model_csr.py:178-192 with UnitMatrixK/M) ---------
 * Row-owner, atomic-free: each output block is summed in a fixed order
 * (a lane-shuffle butterfly over its tasks) and written exactly once.  young: per-element modulus (E)
 * or NULL to use young_scalar.  k_vals (nullable): 9*nnzb values of dtype
 * k_dtype in k_layout.  m_blocks (nullable): nnzb f64 scalars, the mass
 * coupling of each node pair (the value the reference replicates on the
 * three DOF diagonals).  Needs a mesh with row groups. */
int fem_hex8_assemble(const fem_mesh* mesh, const double* young, double young_scalar,
                      double nu, double rho, void* k_vals, int k_dtype, int k_layout,
                      double* m_blocks, void* stream);

/* LDS bytes one work-group of fem_hex8_assemble uses for a row group of the
 * given sizes (to check a grouping fits the 160 KB of a CU before launch). */
int64_t fem_hex8_assemble_lds_bytes(int32_t max_group_elems, int32_t max_group_rows, int k_dtype,
                                    int want_k, int want_m);

/* ---- large displacement (replaces UnitK / InstallMatrixKN / LDinit,
 *      FEM_model_bsr.py:178-191, 436-449, 476-598, behind Get_MatrixKN :134-145)
 * At the displacement u (3N, device), writes the tangent KT and the matrix KM
 * of the residual KM u - F (KM is not symmetric) into kt_vals / km_vals (9*nnzb
 * f64 each, k_layout).  The geometric term uses the element stress summed
 * over the Gauss points visited so far, as the reference's element field S
 * is.  Needs row groups of at most 32 elements (the fem_mesh given is one
 * built for this kernel's LDS budget: fem_hex8_kn_lds_bytes). */
int fem_hex8_assemble_kn(const fem_mesh* mesh, const double* young, double young_scalar, double nu,
                         const double* u, double* kt_vals, double* km_vals, int k_layout, void* stream);
/* K_L + K_sigma(u): the linear stiffness plus the stress-stiffening
 * (geometric) matrix of the linear stress of u, K_sigma_jk = sum_g detJ
 * (a_j^T sigma_g a_k) I with sigma_g = D sym(du/dx) at each Gauss point.  No
 * counterpart in the reference (its Readme.md:189 lists prestressed modal
 * analysis as missing); same row groups and limits as fem_hex8_assemble_kn,
 * whose kernel it shares (F = I, no stress accumulation). */
int fem_hex8_assemble_ks(const fem_mesh* mesh, const double* young, double young_scalar, double nu,
                         const double* u, double* k_vals, int k_layout, void* stream);
/* Spin softening: k_vals -= m_jk (|w|^2 I - w w^T) per node-pair block, m_jk
 * the consistent-mass scalars of fem_hex8_assemble (m_blocks), w = omega3
 * (host pointer, 3 doubles, rad/s).  No counterpart in the reference
 * (Readme.md:189). */
int fem_spin_softening(const fem_mesh* mesh, const double* m_blocks, const double* omega3,
                       double* k_vals, int k_layout, void* stream);
int64_t fem_hex8_kn_lds_bytes(int32_t max_group_elems, int32_t max_group_rows);

/* ---- a8: penalty (replaces ModifyK FEM_model_csr.py:233-242 and
 *      ModifyMatrixK FEM_model_bsr.py:248-259) ------------------------------
 * For each bound entry b (node index bound_node[b], resolved by the host
 * through Nind exactly as the reference does) and DOF j with
 * bound_flag[3*b+j] != -1, SETS the node's diagonal value (j,j) to `value`
 * (the reference uses 1e20).  Entries whose node is outside [0, N) or whose
 * row has no diagonal block are skipped, as the reference's search skips
 * them. */
int fem_apply_penalty(const fem_mesh* mesh, const int32_t* bound_node, const double* bound_flag,
                      int64_t num_bounds, void* k_vals, int k_dtype, int k_layout, double value,
                      void* stream);

/* ---- a9: load vector (replaces InstallVectorF FEM_model_csr.py:214-229 and
 *      LoadF :205-210) -------------------------------------------------------
 * f_out (3N) = centrifugal body force of every element gathered per node in
 * element order + nodal_load (3N, nullable).  work: 24*E doubles. */
/* Row-owner kernels on the mesh's row groups (one work-group per group, the
 * group's element values kept in LDS); they need grp_row_rec and no scratch:
 * fem_body_force_work_bytes / fem_stress_work_bytes return 0 (the work
 * arguments are kept for the ABI and ignored).  LDS: 200 B (load vector) or
 * 392 B (stress) per element of the largest group, at most 160 KB. */
int64_t fem_body_force_work_bytes(int64_t num_elem);
int fem_body_force(const fem_mesh* mesh, double rho, const double* omega3,
                   const double* nodal_load, double* f_out, void* work, int64_t work_bytes,
                   void* stream);

/* ---- a13: stress recovery (replaces get_NdSol/get_NdAve/UnitVectorS
 *      POST_process.py:172-209 and the host eigh/von Mises :79-95) ---------
 * u (3N) displacements; out (N x 10): [sxx syy szz txy tyz txz, p1 p2 p3
 * (descending), the reference's von Mises term]. */
int64_t fem_stress_work_bytes(int64_t num_elem);
int fem_stress_recover(const fem_mesh* mesh, const double* young, double young_scalar,
                       double nu, const double* u, double* out, void* work,
                       int64_t work_bytes, void* stream);

/* ---- SpMV / SpMM on the BSR(3x3) pattern (the K/M products inside
 *      Lanczos_solve SOLVE_eigen.py:65-74 and the PCG forms) -------------- */

/* y = A x for the 3x3-block matrix (sm_node, re_node, vals); dtype F64,
 * F32 (x, y of the value dtype) or F32F64 (f32 values, f64 x and y). */
int fem_bsr_spmv(const fem_mesh* mesh, const void* vals, int dtype, const void* x, void* y,
                 void* stream);
/* Y = A X for nrhs column-interleaved right-hand sides X[(3*i+p)*nrhs + c]. */
int fem_bsr_spmm(const fem_mesh* mesh, const double* vals, const double* X, double* Y,
                 int64_t nrhs, void* stream);
/* fem_bsr_spmm with 16 columns on a structured node grid of nx x ny x nz
 * nodes (node id = ix + nx*(iy + ny*iz), the pattern's columns within one
 * index step of the row along each axis, as for any hex8 grid of that
 * numbering): a work-group walks up node planes of an 8 x 4 tile and keeps
 * three planes of X in LDS, so X is fetched once per plane instead of
 * gathered per block.  Deterministic; the sums run in a different order than
 * fem_bsr_spmm's, so results agree to rounding.  X 16-byte aligned, Y != X. */
int fem_bsr_spmm16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const double* vals,
                        const double* X, double* Y, void* stream);
/* fem_bsr_spmm16_grid that also leaves, per work-group w and column c,
 * part[w*16 + c] = sum over the work-group's rows i of X_ic Y_ic (the p.q of
 * flexible CG read off the product; no counterpart in the reference, whose
 * solves are direct).  part holds part_len doubles, at least
 * 16 * fem_spmm16_grid_partials(nx, ny, nz); the count is per device (it
 * follows the CU count), so query it on the device that runs the product. */
int64_t fem_spmm16_grid_partials(int64_t nx, int64_t ny, int64_t nz);
int fem_bsr_spmm16_grid_pq(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const double* vals,
                           const double* X, double* Y, double* part, int64_t part_len, void* stream);
/* fem_mass_spmm with 16 columns on a structured node grid (the layout and
 * conditions of fem_bsr_spmm16_grid). */
int fem_mass_spmm16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const double* m_blocks,
                         const double* X, double* Y, void* stream);
/* y = M x with the scalar-per-block mass (the reference's CSR M). */
int fem_mass_spmm(const fem_mesh* mesh, const double* m_blocks, const double* X, double* Y,
                  int64_t nrhs, void* stream);

/* ---- PCG building blocks (new GPU solver forms replacing the Pardiso call
 *      SOLVE_static.py:31-38) ------------------------------------------------ */

/* dinv (N x 9, dtype) = inverse of each node's 3x3 diagonal block.
 * FEM_E_NUMERIC is reported through the flag (device int) when a block is
 * singular or a row has no diagonal block. */
int fem_bj_setup(const fem_mesh* mesh, const void* vals, int dtype, void* dinv,
                 int32_t* flag, void* stream);

/* Scratch bytes of the fused PCG kernels for n3 = 3N unknowns. */
int64_t fem_pcg_work_bytes(int64_t num_node);

/* The PCG scalar state lives in device memory (8 doubles): state[0] = rz
 * used by the current iteration, [1] = rz of the latest residual, [2] =
 * alpha, [3] = p.q, [4] = r.r, [5] = b.b, [6] = beta. */
/* r = b - A x; z = Dinv r; p = z; rz, rr, bb. */
int fem_pcg_start(const fem_mesh* mesh, const void* vals, const void* dinv, int dtype,
                  const void* b, void* x, void* r, void* z, void* p, double* state,
                  void* work, int64_t work_bytes, void* stream);
/* One iteration: q = A p; alpha = rz / p.q; x += alpha p; r -= alpha q;
 * z = Dinv r; beta = r.z / rz; p = z + beta p; rr = r.r.  Dot products are
 * reduced in a fixed order without atomics (bit-reproducible). */
int fem_pcg_iterate(const fem_mesh* mesh, const void* vals, const void* dinv, int dtype,
                    void* x, void* r, void* z, void* p, void* q, double* state,
                    void* work, int64_t work_bytes, void* stream);
/* Residual replacement (FP64 vectors): r = b - y where y = A64 x was
 * computed by the caller, z = Dinv r, rz and rr refreshed in state; when
 * p_restart is non-NULL the search direction restarts, p = z. */
int fem_pcg_replace_residual(const fem_mesh* mesh, const void* dinv, int dtype, const void* b,
                             const void* y, void* r, void* z, void* p_restart, double* state,
                             void* work, int64_t work_bytes, void* stream);

/* ---- multi-RHS block-Jacobi PCG: the shift-invert solves K^-1 (M u) of
 *      block Lanczos (SOLVE_eigen.py:65, there a Pardiso solve) ------------
 * Blocks are node-interleaved with columns fastest: X[(3*i+p)*nrhs + c];
 * nrhs is a power of two <= 64 (pad with zero columns).  state is 8 x nrhs
 * doubles (rows: rz, rz_next, alpha, p.q, r.r, b.b, beta, spare). */
int64_t fem_mpcg_work_bytes(int64_t nrhs);
int fem_mpcg_start(const fem_mesh* mesh, const double* vals, const double* dinv, int64_t nrhs,
                   const double* B, double* X, double* R, double* Z, double* P, double* state,
                   void* work, int64_t work_bytes, void* stream);
int fem_mpcg_iterate(const fem_mesh* mesh, const double* vals, const double* dinv, int64_t nrhs,
                     double* X, double* R, double* Z, double* P, double* Q, double* state,
                     void* work, int64_t work_bytes, void* stream);

/* ---- geometric multigrid V-cycle on structured node grids (FP32), the
 *      preconditioner of the FP64 flexible CG static solve that replaces the
 *      Pardiso call of SOLVE_static.py:31-38; no counterpart in the
 *      reference) -------------------------------------------------------------
 * Node grids are numbered x fastest, n = ix + nf[0]*(iy + nf[1]*iz).  The
 * coarse grid keeps every second index along each axis plus the last one;
 * prolongation P is trilinear in the index coordinate, restriction is P^T.
 * Per axis a (tables concatenated over a = 0, 1, 2):
 *   left[f]    coarse index at or left of fine index f,
 *   wright[f]  weight of coarse index left[f] + 1 (0 when f coincides with
 *              a coarse node; then left[f] + 1 is never read),
 *   span[2c], span[2c+1]  first and last fine index with a nonzero weight
 *              for coarse index c. */
typedef struct fem_grid_transfer {
  int32_t nf[3];
  int32_t nc[3];
  const int32_t* left;    /* nf[0] + nf[1] + nf[2] */
  const float* wright;    /* nf[0] + nf[1] + nf[2] */
  const int32_t* span;    /* 2 * (nc[0] + nc[1] + nc[2]) */
} fem_grid_transfer;

/* FP32 block-Jacobi smoother pieces on the FP32 matrix vals, in layout
 * FEM_LAYOUT_BSR or FEM_LAYOUT_CSR (the value orders of fem_hex8_assemble):
 * mode 0: out = b - K x;  mode 1: out = x + omega Dinv (b - K x) (out must
 * not alias x);  mode 2: out = omega Dinv b (vals and x unused). */
int fem_mg_smooth(const fem_mesh* mesh, const float* vals, int layout, const float* dinv, float omega,
                  const float* x, const float* b, float* out, int mode, void* stream);
/* coarse = P^T fine, for nrhs column-interleaved vectors (3 FP32 values per
 * node and column: v[(3*i + p)*nrhs + c]; nrhs in [1, 64]). */
int fem_mg_restrict(const fem_grid_transfer* t, const float* fine, float* coarse, int64_t nrhs,
                    void* stream);
/* fine += P coarse (same layout). */
int fem_mg_prolong_add(const fem_grid_transfer* t, const float* coarse, float* fine, int64_t nrhs,
                       void* stream);
/* y = A x, A dense n x n FP32 row-major (the coarsest level's inverse). */
int fem_dense_gemv_f32(const float* A, int64_t n, const float* x, float* y, void* stream);
/* The smoother modes of fem_mg_smooth (BSR values) on nrhs column-interleaved
 * vectors at once (the inner solves of block Lanczos). */
int fem_mg_smooth_multi(const fem_mesh* mesh, const float* vals, const float* dinv, float omega,
                        const float* X, const float* B, float* OUT, int64_t nrhs, int mode, void* stream);
/* fem_mg_smooth_multi's residual (mode 0) and update (mode 1) passes for 16
 * columns on a structured node grid (the layout and conditions of
 * fem_bsr_spmm16_grid). */
int fem_mg_smooth16_grid(const fem_mesh* mesh, int64_t nx, int64_t ny, int64_t nz, const float* vals,
                         const float* dinv, float omega, const float* X, const float* B, float* OUT, int mode,
                         void* stream);
/* Y = A X, A dense n x n FP32 row-major, X and Y n x nrhs row-major. */
int fem_dense_gemm_nb_f32(const float* A, int64_t n, const float* X, float* Y, int64_t nrhs, void* stream);

/* C (p x q, row-major) = A^T B for tall A (n x p, row stride lda) and B
 * (n x q, row stride ldb), FP64: the inner products of block Lanczos
 * (SOLVE_eigen.py:56, 67, 72-73, 76).  Reduced over row chunks into
 * partials in work (fem_gemm_tn_work_bytes), then summed in a fixed order
 * (deterministic).  p <= 512, q <= 64, p*q <= 8192. */
int64_t fem_gemm_tn_work_bytes(int64_t p, int64_t q);
int fem_gemm_tn_f64(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t n, int64_t p, int64_t q,
                    double* C, void* work, int64_t work_bytes, void* stream);

/* fem_gemm_tn_f64 without a workspace, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The same kernels
 * and launch shape as fem_gemm_tn_f64, so C is bit-identical to it; the
 * partials live in a static device array of FEM_GEMM_TN_NOALLOC_MAX_PQ x 1024
 * doubles (16 MiB), held by each GPU the library runs on until the process
 * ends, and its address is looked up on the first call on each device.  The
 * call allocates nothing.  p * q above FEM_GEMM_TN_NOALLOC_MAX_PQ gives
 * FEM_E_LIMIT.  Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams, other host threads) race and must be serialised by
 * the caller; consecutive calls on one stream are safe (each call writes every
 * partial it reads). */
#define FEM_GEMM_TN_NOALLOC_MAX_PQ 2048
int fem_gemm_tn_f64_noalloc(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t n, int64_t p,
                            int64_t q, double* C, void* stream);

/* Y = alpha * A C + beta * Y for a tall A (n x k, row stride lda) and a small
 * C (k x q, row stride ldc), FP64, Y (n x q, row stride ldy) not overlapping
 * A: the block updates of block Lanczos (SOLVE_eigen.py:55 u <- u R^-1, :60
 * v -= u T_ii, :66-71 the two reorthogonalisation passes v -= Q (Q^T M v)).
 * Each output is summed in ascending column order (deterministic).
 * q <= 64; beta == 0 does not read Y. */
int fem_gemm_nn_f64(const double* A, int64_t lda, const double* C, int64_t ldc, double* Y, int64_t ldy,
                    int64_t n, int64_t k, int64_t q, double alpha, double beta, void* stream);

/* Upper Cholesky factor R (A = R^T R, R upper triangular, zeros below) of a
 * small SPD b x b matrix A (row-major, device) and its inverse Rinv, both b x b
 * row-major: the scipy.linalg.cholesky and the inverse that scale each new
 * Lanczos block (SOLVE_eigen.py:52-53, 75-77).  b <= 64, one work-group.  A
 * non-positive pivot makes R and Rinv NaN from that row on (no host sync;
 * the caller checks the result). */
int fem_chol_upper_inv_f64(const double* A, int64_t b, double* R, double* Rinv, void* stream);

/* Flexible CG, FCG(1), vector steps on nb column-interleaved vectors (n rows,
 * entry (i, c) at i*nb + c), around an outside preconditioner (the multigrid
 * V-cycle).  The reference's static and shift-invert solves are direct
 * (SOLVE_static.py:12-38, SOLVE_eigen.py:51); this is the iterative solver
 * that replaces them.  mask (n doubles, 1 on free DOFs, 0 on penalised ones)
 * may be NULL (all free).  work (fem_fcg_work_bytes) holds the deterministic
 * dot-product partials and the per-column scalars that carry from one
 * fem_fcg_update to the next fem_fcg_direction: pass the same buffer.
 *   direction: P = mask*Z - beta*P, beta = (mask Z . Qp) / (previous p.q);
 *              first != 0: P = mask*Z.  Z is FEM_DTYPE_F32 or _F64.
 *   update:    alpha = p.r / p.q on columns with active[c] != 0 (NULL: all),
 *              0 elsewhere; X += alpha P; R -= alpha mask*Q; R32 = (float)R
 *              (R32 may be NULL); rr[c] = |R_c|^2 (device array of nb).
 * nb <= 64. */
int64_t fem_fcg_work_bytes(int64_t nb);
int fem_fcg_direction(int64_t n, int64_t nb, const void* Z, int z_dtype, const double* mask, const double* Qp,
                      double* P, int first, void* work, int64_t work_bytes, void* stream);
int fem_fcg_update(int64_t n, int64_t nb, const double* P, const double* Q, const double* mask,
                   const int32_t* active, double* X, double* R, float* R32, double* rr, void* work,
                   int64_t work_bytes, void* stream);
/* The same two steps with p.q taken from fem_bsr_spmm16_grid_pq's partials
 * instead of a pass over P and Q (no counterpart in the reference):
 *   direction_pr: fem_fcg_direction that also leaves the partials of p.r of
 *                 the new P (R: the current residual) in work;
 *   update_split: alpha = p.r / p.q from those and the pq_blocks partials of
 *                 pq_part (pq_part[w*nb + c]), then fem_fcg_update's X, R,
 *                 R32 and rr.
 * Between the two, run the product Q = K P that fills pq_part. */
int fem_fcg_direction_pr(int64_t n, int64_t nb, const void* Z, int z_dtype, const double* mask,
                         const double* Qp, const double* R, double* P, int first, void* work, int64_t work_bytes,
                         void* stream);
int fem_fcg_update_split(int64_t n, int64_t nb, const double* P, const double* Q, const double* mask,
                         const int32_t* active, const double* pq_part, int64_t pq_blocks, double* X, double* R,
                         float* R32, double* rr, void* work, int64_t work_bytes, void* stream);
/* The residual of the block solve (multigrid.py solve_block_fcg_mg's start,
 * true-residual checks and final norms, in one pass; no counterpart kernel in
 * the reference): t = (B - Y) * mask (Y NULL: B * mask; mask NULL: all
 * ones); out[c] = |t_c|^2 and out[nb + c] = |B_c|^2 (device array of 2 nb);
 * columns with sel[c] != 0 (sel NULL: all) also get R = t and R32 =
 * (float)t (either may be NULL).  Uses the work buffer's partials only, so
 * the scalars fem_fcg_update carries to fem_fcg_direction survive. */
int fem_fcg_residual(int64_t n, int64_t nb, const double* B, const double* Y, const double* mask,
                     const int32_t* sel, double* R, float* R32, double* out, void* work, int64_t work_bytes,
                     void* stream);

/* ---- 3x3x3 mesh convolution on the structured node grid (the learned
 *      preconditioner; no counterpart in the reference) ---------------------
 * BF16 operands on MFMA, FP32 accumulation.  Activations channels-last
 * x[p*C + c], p = ix + nx*(iy + ny*iz); cin a multiple of 32 (<= 64), cout a
 * multiple of 16 (<= 64); weights w[tap][co][ci], tap = (dz+1)*9 +
 * (dy+1)*3 + (dx+1); zero padding outside the grid.  Pointers to bf16 data
 * are void*. */
/* y = conv(x) + bias; writes any of y_act (bf16, ReLU when relu != 0),
 * y_pre (bf16, before ReLU), y_f32 (f32, before ReLU). */
int fem_meshconv_fwd_bf16(const void* x, int cin, const void* w, const float* bias, int cout, int nx,
                          int ny, int nz, void* y_act, void* y_pre, float* y_f32, int relu,
                          void* stream);
/* wt[tap][ci][co] = w[26-tap][co][ci], padded to (cout_t, cin_t): the
 * weights of the backward-data convolution. */
int fem_meshconv_flip_bf16(const void* w, int cout, int cin, void* wt, int cout_t, int cin_t,
                           void* stream);
/* dw[tap][co][ci] (f32) = sum_p dy[p][co] x[p + off(tap)][ci], split over
 * pixel chunks and reduced in chunk order. */
int64_t fem_meshconv_wgrad_work_bytes(int cin, int cout, int nx, int ny, int nz, int chunk);
int fem_meshconv_wgrad_bf16(const void* x, int cin, const void* dy, int cout, int nx, int ny, int nz,
                            int chunk, float* dw, void* work, int64_t work_bytes, void* stream);
/* out[c] = sum_p a[p*C + c] (a f32 or bf16), deterministic. */
int64_t fem_colsum_work_bytes(int64_t P, int C, int chunk);
int fem_colsum(const void* a, int is_bf16, int64_t P, int C, int chunk, float* out, void* work,
               int64_t work_bytes, void* stream);
/* dz = da * (z > 0), bf16. */
int fem_relu_bwd_bf16(const void* da, const void* z, int64_t n, void* dz, void* stream);
/* out (P x cd, bf16) = a (P x cs, f32), channels cs..cd-1 zero. */
int fem_to_bf16(const float* a, int64_t P, int cs, int cd, void* out, void* stream);
/* The first layer with 4 input channels, packed along K (k = 4*tap + ci, 4
 * MFMA k-steps per tile instead of 27): x4 (P x 4 bf16, 8-byte aligned), wp
 * from fem_meshconv_pack_c4_bf16 (16-byte aligned); writes y_act (ReLU when
 * relu != 0) and/or y_pre, bf16.  Same function as fem_meshconv_fwd_bf16 on
 * the input padded to 32 channels, to FP32 summation order. */
int fem_meshconv_fwd_c4_bf16(const void* x4, const void* wp, const float* bias, int cout, int nx, int ny,
                             int nz, void* y_act, void* y_pre, int relu, void* stream);
/* wp[s][co][kk] (4 x cout x 32 bf16) = w[tap][co][ci] at 32*s + kk = 4*tap + ci
 * (tap < 27, ci < 4; zero elsewhere), from w[27][cout][cin_w]. */
int fem_meshconv_pack_c4_bf16(const void* w, int cin_w, int cout, void* wp, void* stream);
/* x4[p][c] = bf16(float(r[3p+c] / *scale)) for c < 3 and bf16(efeat[p]) for
 * c = 3: the network input from a residual r (3P doubles) and the per-node
 * feature efeat (P floats); scale is a device pointer to one double. */
int fem_meshconv_features_c4(const double* r, int64_t P, const double* scale, const float* efeat,
                             void* x4, void* stream);
/* c[3p+j] = (double)y[p*ld + j] * (*scale / e_ref) * mask[3p+j] (mask NULL: 1),
 * j < 3: the correction from the network output y (P x ld floats). */
int fem_meshconv_correction(const float* y, int ld, int64_t P, const double* scale, double e_ref,
                            const double* mask, double* c, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FEM355_H */
