/*
 * mpk_assembly.h -- C ABI of the MI355X element-assembly engine for MagPackage.
 *
 * One shared library, libmpk_assembly.so (gfx950 HIP kernels).  Plain
 * pointers and sizes only.  Every function returns an mpk status code
 * (MPK_OK == 0); on failure mpk_last_error() returns a thread-local message.
 *
 * Pointer spaces:
 *   - "device" arrays live in HBM of the current HIP device;
 *   - "host" arrays live in host memory;
 *   - quadrature tables are always host arrays (they are copied into kernel
 *     arguments).
 * Streams: `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Calls are asynchronous on `stream` unless stated otherwise; device outputs
 * are valid once the stream has run them.  The library never allocates the
 * caller's outputs.
 *
 * Index conventions follow the reference (MATLAB): node ids, DOF ids and
 * region tags are 1-based; a DOF of 0 marks an inactive entity, whose
 * triplets Assembler.m filters out (src/Assembler.m:37).
 */
#ifndef MPK_ASSEMBLY_H
#define MPK_ASSEMBLY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define MPK_OK 0
#define MPK_ERR_ARGS 1   /* wrong argument count/shape/null pointer ("MagPackage:Args") */
#define MPK_ERR_INDEX 2  /* node / DOF / tag index outside its range */
#define MPK_ERR_HIP 3    /* HIP runtime error */
#define MPK_ERR_LIMIT 4  /* a size limit of the engine (message says which) */

/* ---- precision modes of the CSC assembly --------------------------------- */
/* FP64, reference operation order, no FMA contraction: bit-identical to the
 * MEX kernels followed by sparse() summing duplicates in triplet order. */
#define MPK_PREC_FP64_EXACT 0
/* FP64 closed forms with FMA (default): <= 1e-12 relative to FP64_EXACT. */
#define MPK_PREC_FP64 1
/* Local block entries in FP32, determinant/inverse and CSC accumulation in
 * FP64: <= 1e-6 relative. */
#define MPK_PREC_FP32_LOCAL 2

int mpk_abi_version(void);
const char* mpk_last_error(void);
/* Number of visible HIP devices (0 when none). */
int mpk_device_count(void);

/* Quadrature tables in the reference's layouts (host memory).
 * qw[nq]; ref_curl and ref_val are [6 x 3 x nq] column-major (index
 * i + 6*c + 18*q), as produced by nedelec_tet_p1 (src/shape_functions/
 * nedelec_tet_p1.m:18-51) for get_quadrature_data (src/utils/
 * get_quadrature_data.m:12-29); q_pts is [3 x nq] column-major.  nq <= 8. */
typedef struct mpk_quadrature {
    int32_t nq;
    double qw[8];
    double q_pts[24];
    double ref_val[144];
    double ref_curl[144];
} mpk_quadrature;

/* Packed B-H material tables, the 8 SoA arrays Assembler.packMaterialData
 * builds (src/Assembler.m:270-304) and assemble_jacobian_kernel_mex reads
 * (src/mex/MexMaterialUtils.hpp:9-18).  Device arrays. */
typedef struct mpk_material {
    const double* linear_nu;    /* [num_tags] */
    const double* is_nonlinear; /* [num_tags] */
    const double* max_bsq;      /* [num_tags] */
    const double* breaks;       /* [num_breaks] */
    const double* coefs;        /* [num_coefs], per piece c3 c2 c1 c0 */
    const double* spline_start; /* [num_tags], 0-based offsets into breaks */
    const double* spline_count; /* [num_tags], number of pieces */
    const double* coef_start;   /* [num_tags], 0-based offsets into coefs */
    int64_t num_tags;
    int64_t num_breaks;
    int64_t num_coefs;
} mpk_material;

/* ========================================================================== */
/* 1. MEX-compatible element kernels: COO triplets, reference layouts          */
/* ========================================================================== */
/* All arrays device, doubles, MATLAB layouts: P [3 x nn], T [4 x ne] (1-based),
 * cell_dofs / signs [6 x ne], per-element coefficients [ne].  Outputs are
 * 36*ne triplets in the reference order: slot 36*e + 6*c + r holds
 * (dofs[r], dofs[c], K[r][c]*s[r]*s[c]).  Node ids are checked against nn
 * (MPK_ERR_INDEX); these calls synchronise `stream` for that check. */

/* Replaces assemble_curl_curl_kernel_mex (src/mex/assemble_curl_curl_kernel_mex.cpp:18-121). */
int mpk_curl_curl_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                      const double* cell_dofs, const double* signs, const double* nu,
                      const mpk_quadrature* quad, double* I_out, double* J_out,
                      double* V_out, void* stream);

/* Replaces assemble_mass_kernel_mex (src/mex/assemble_mass_kernel_mex.cpp:16-127).
 * coeffs may be NULL or have num_coeffs != ne: sigma = 1 then (:34,66). */
int mpk_mass_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                 const double* cell_dofs, const double* signs, const double* coeffs,
                 int64_t num_coeffs, const mpk_quadrature* quad, double* I_out,
                 double* J_out, double* V_out, void* stream);

/* Replaces assemble_jacobian_kernel_mex (src/mex/assemble_jacobian_kernel_mex.cpp:7-153).
 * sol_a has num_sol entries (indexed dof-1).  calc_j == 0: I/J/V untouched
 * (the MEX returns 0x1 arrays), R triplets still written. */
int mpk_jacobian_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                     const double* cell_dofs, const double* signs, const double* tags,
                     const double* sol_a, int64_t num_sol, const mpk_quadrature* quad,
                     const mpk_material* mat, int calc_j, double* I_out, double* J_out,
                     double* V_out, double* R_idx, double* R_val, void* stream);

/* Replaces assemble_source_kernel_mex (src/mex/assemble_source_kernel_mex.cpp:16-120).
 * Js [3 x ne] device current density per element; quad carries Qw and
 * ValRef.  Outputs 6*ne pairs, slot 6*e + i = (dofs[i], F_i * s[i]).
 * Synchronises `stream` for the node-index check. */
int mpk_source_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                   const double* cell_dofs, const double* signs, const double* Js,
                   const mpk_quadrature* quad, double* I_out, double* V_out, void* stream);

/* Replaces assemble_scalar_laplacian_kernel_mex
 * (src/mex/assemble_scalar_laplacian_kernel_mex.cpp:13-105):
 * K = (sigma grad phi_i, grad phi_j) over P1 nodal functions.  node_dofs
 * [4 x ne] (1-based, doubles), coeffs [ne] sigma (|sigma| <= 1e-12: zero
 * block, indices still written); quad carries Qw (its tables are unused);
 * grad_ref [4 x 3 x nq] host array at i + 4c + 12q.  Outputs 16*ne triplets,
 * slot 16*e + 4*c + r = (dofs[r], dofs[c], K[r][c]).  Device arrays except
 * quad and grad_ref.  Synchronises `stream` for the node-index check. */
int mpk_scalar_laplacian_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                             const double* node_dofs, const double* coeffs,
                             const mpk_quadrature* quad, const double* grad_ref,
                             double* I_out, double* J_out, double* V_out, void* stream);

/* Replaces assemble_coupling_kernel_mex (src/mex/assemble_coupling_kernel_mex.cpp:13-131):
 * the A-V coupling block C = (sigma N_i, grad phi_j), 6 edge rows x 4 node
 * columns.  edge_dofs / signs [6 x ne], node_dofs [4 x ne], coeffs [ne];
 * quad->ref_val is the Nedelec table NedRef, grad_ref the [4 x 3 x nq] host
 * table.  Outputs 24*ne triplets, slot 24*e + 6*c + r =
 * (edge_dofs[r], node_dofs[c], C[r][c] * signs[r]).  Synchronises `stream`. */
int mpk_coupling_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                     const double* edge_dofs, const double* signs, const double* node_dofs,
                     const double* coeffs, const mpk_quadrature* quad, const double* grad_ref,
                     double* I_out, double* J_out, double* V_out, void* stream);

/* Replaces assemble_winding_kernel_mex (src/mex/assemble_winding_kernel_mex.cpp:13-126):
 * F_i = (scale * J_dir, N_i) over the elements with (int)tags[e] == (int)coil_id;
 * other elements give (0, 0) pairs.  J_dir from dir_field [3 x ne] (device, may be
 * NULL) else from uniform_dir (host, 3 doubles).  quad carries Qw and ValRef.
 * Outputs 6*ne pairs, slot 6*e + i = (dofs[i], F_i * signs[i]).  Synchronises
 * `stream` for the node-index check. */
int mpk_winding_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                    const double* cell_dofs, const double* signs, const double* tags,
                    double coil_id, const double* dir_field, const double* uniform_dir,
                    double scale, const mpk_quadrature* quad, double* I_out, double* V_out,
                    void* stream);

/* Replaces assemble_hbfem_kernel_mex (src/mex/assemble_hbfem_kernel_mex.cpp:25-260):
 * harmonic-balance residual and DC-reluctivity block Jacobian.  sol_r / sol_i
 * [num_dofs x num_harm] (sol_i may be NULL), dmat_r / dmat_i [num_time x num_harm]
 * (AFT D matrix; dmat_i may be NULL), pmat_r / pmat_i [num_harm x num_time] (AFT P
 * matrix; pmat_i may be NULL), column-major; scalings [num_harm] (Jacobian block
 * factors).  The material table needs coef_start (the MEX leaves it
 * uninitialised).  Outputs, in the MEX's slots: R rows / cols / real / imag at
 * 6*H*e + 6*h + k; with request_j, 36*H*e triplets (dofs + h*num_dofs) at
 * 36*H*e + 36*h + 6*c + r.  A DOF of 0 reads sol at h*num_dofs - 1 for h >= 1 (the
 * MEX's index arithmetic) and 0 for h = 0.  num_harm <= 64.  Device arrays except
 * quad and mat; synchronises `stream` for the index checks. */
int mpk_hbfem_coo(const double* P, int64_t nn, const double* T, int64_t ne,
                  const double* cell_dofs, const double* signs, const double* tags,
                  const double* sol_r, const double* sol_i, int64_t num_dofs, int32_t num_harm,
                  const double* dmat_r, const double* dmat_i, int32_t num_time,
                  const double* pmat_r, const double* pmat_i, const mpk_quadrature* quad,
                  const mpk_material* mat, int request_j, const double* scalings,
                  double* I_out, double* J_out, double* V_out, double* R_rows, double* R_cols,
                  double* R_val_r, double* R_val_i, void* stream);

/* PostProcessor.computeElementB (src/post/PostProcessor.m:64-126) for one
 * solution column: B_out [3 x ne] device, B_e = curl A at the element.  T
 * [ne x 4] and cell_dofs [ne x 6] device int32, 1-based (0 = inactive DOF);
 * signs from the node ids (Mesh.T2E_Sign); quad->ref_curl row q = 0 is the
 * (constant) reference curl table.  Synchronises `stream` for index checks. */
int mpk_element_b(const double* P, int64_t nn, const int32_t* T, const int32_t* cell_dofs,
                  int64_t ne, const double* sol, int64_t num_sol, const mpk_quadrature* quad,
                  double* B_out, void* stream);

/* evaluate_nu_derivative (src/mex/MexMaterialUtils.hpp:24-87) at n points:
 * s [n] (|B|^2) and tags [n] (1-based int32) device; nu_out [n], dnu_out [n]
 * (may be NULL).  Synchronises `stream` for the tag check. */
int mpk_eval_nu(const double* s, const int32_t* tags, int64_t n, const mpk_material* mat,
                double* nu_out, double* dnu_out, void* stream);

/* ========================================================================== */
/* 2. COO -> CSC with MATLAB sparse() semantics                               */
/* ========================================================================== */
/* Replaces `valid = (I>0)&(J>0); sparse(I(valid),J(valid),V(valid),m,n)`
 * (src/Assembler.m:37-38).  Device arrays.  Duplicates are summed in triplet
 * order; exact zeros are dropped when drop_zeros != 0 (MATLAB's rule).
 * colptr_out [n+1] (int64), rowind_out [cap] (0-based int32), vals_out [cap].
 * *nnz_out receives nnz; if nnz > cap, MPK_ERR_LIMIT is returned and the
 * outputs are not written (cap = nt always suffices).  Synchronises `stream`. */
int mpk_sparse_coo_to_csc(const double* I, const double* J, const double* V, int64_t nt,
                          int64_t m, int64_t n, int drop_zeros, int64_t* colptr_out,
                          int32_t* rowind_out, double* vals_out, int64_t cap,
                          int64_t* nnz_out, void* stream);

/* ========================================================================== */
/* 3. Mesh topology and DOF map (device, engine-native int32 layouts)          */
/* ========================================================================== */
/* Replaces Mesh.generateEdges (src/Mesh.m:114-138).  T: [ne x 4] rows of
 * 1-based node ids (== MATLAB T [4 x ne] column-major).  edges_out: [2 x cap]
 * (cap >= 6*ne) 1-based sorted unique (min,max) pairs; t2e_out [6 x ne]
 * 1-based edge ids; sign_out [6 x ne] +1 iff T(a) <= T(b).  Synchronises. */
int mpk_generate_edges(const int32_t* T, int64_t ne, int64_t nn, int32_t* edges_out,
                       int32_t* t2e_out, int8_t* sign_out, int64_t* nedge_out,
                       void* stream);

/* Replaces DofHandler.distributeDofs, Nedelec branch (src/DofHandler.m:72-105).
 * active_elem [ne] (NULL = all active); entity_map [nedge]; cell_dofs [6 x ne].
 * Synchronises. */
int mpk_distribute_dofs(const int32_t* t2e, int64_t ne, int64_t nedge,
                        const uint8_t* active_elem, int64_t offset, int32_t* entity_map,
                        int32_t* cell_dofs, int64_t* ndofs_out, void* stream);

/* ========================================================================== */
/* 4. CSC assembly plan: pattern + incidence lists, built once per mesh/DOF map */
/* ========================================================================== */
/* The pattern is sparse()'s structural pattern of the 36*ne triplets with
 * both DOFs > 0 (no value-dependent dropping; see mpk_csc_drop_zeros).  For
 * every column j it stores the incident (element, local edge) pairs in
 * ascending element order, so accumulation follows triplet order. */
typedef struct mpk_plan mpk_plan;

/* T: [ne x 4] 1-based node ids, cell_dofs: [ne x 6] 1-based DOFs (0 =
 * inactive), both device int32.  ndofs = NumGlobalDofs.  Synchronises. */
int mpk_plan_create(const int32_t* T, const int32_t* cell_dofs, int64_t ne, int64_t nn,
                    int64_t ndofs, void* stream, mpk_plan** plan_out);
int mpk_plan_destroy(mpk_plan* plan);
/* nnz, number of columns, elements, incidences and the largest LDS tile. */
int mpk_plan_info(const mpk_plan* plan, int64_t* nnz, int64_t* ncols, int64_t* ne,
                  int64_t* n_incidences, int64_t* max_tile_nnz);
/* *ring != 0 when the closed-form (FP64, FP32_LOCAL) kernels walk each
 * column's elements in rotation order around the edge (every column a simple
 * ring or chain, < 2^24 elements and nodes); else element order.  *ring == 2:
 * the walk reads the compact 8-byte ring records (new node within 2^21 ids
 * of the column's lower node, columns of <= 63 entries); 1: 16-byte ones. */
int mpk_plan_ring_order(const mpk_plan* plan, int* ring);
/* Device pointers owned by the plan: colptr [n+1], rowind [nnz] (0-based). */
int mpk_plan_pattern(const mpk_plan* plan, const int32_t** colptr, const int32_t** rowind);

/* Loads node coordinates P [3 x nn] (device doubles) into the plan's padded
 * copy read by the FP64 / FP32_LOCAL kernels.  Those kernels accept P == NULL
 * and then use this copy: a Newton loop on a fixed mesh loads P once instead
 * of once per assembly.  No reference counterpart (the MEX kernels take P on
 * every call, src/Assembler.m:37-38); FP64_EXACT always needs P. */
int mpk_plan_set_coords(mpk_plan* plan, const double* P, void* stream);

/* CSC values of K (curl-curl, src/mex/assemble_curl_curl_kernel_mex.cpp:18-121
 * + sparse()).  P [3 x nn] device doubles (NULL: coordinates loaded by
 * mpk_plan_set_coords, FP64 / FP32_LOCAL only), nu [ne] device.  vals [nnz]. */
int mpk_csc_curl_curl(const mpk_plan* plan, const double* P, const double* nu,
                      const mpk_quadrature* quad, int precision, double* vals, void* stream);

/* CSC values of the (sigma-)mass (src/mex/assemble_mass_kernel_mex.cpp:16-127
 * + sparse()).  coeffs [ne] device or NULL (sigma = 1). */
int mpk_csc_mass(const mpk_plan* plan, const double* P, const double* coeffs,
                 const mpk_quadrature* quad, int precision, double* vals, void* stream);

/* Complex128 CSC values of K + 1j*omega*M_sigma (src/solver/FrequencySolver.m:
 * 127-131 composition), fused: real = sum K, imag = omega * sum M.
 * vals_c128 [2*nnz] interleaved (re, im).  quad_k / quad_m: stiffness and
 * mass tables (Assembler.m:30-32 and :64-66). */
int mpk_csc_curl_curl_mass_complex(const mpk_plan* plan, const double* P, const double* nu,
                                   const double* sigma, double omega,
                                   const mpk_quadrature* quad_k, const mpk_quadrature* quad_m,
                                   int precision, double* vals_c128, void* stream);

/* Newton Jacobian J and residual R (src/mex/assemble_jacobian_kernel_mex.cpp
 * + sparse() of J and of R, src/Assembler.m:199-216).  tags [ne] device int32
 * (1-based); sol_a [ndofs] device.  j_vals [nnz] (ignored when calc_j == 0);
 * residual [ndofs]. */
int mpk_csc_jacobian(const mpk_plan* plan, const double* P, const int32_t* tags,
                     const double* sol_a, const mpk_material* mat,
                     const mpk_quadrature* quad, int calc_j, int precision,
                     double* j_vals, double* residual, void* stream);

/* Fused Newton set in one pass over the plan (no single reference call: the
 * reference assembles these with three MEX calls + sparse(),
 * src/Assembler.m:26-76,182-231):
 *   k_vals [nnz]  secant stiffness K(nu_e), nu_e = nu(|B_e|^2) from the B-H
 *                 table at the element's B = curl of sol_a (the matrix
 *                 assembleStiffness builds from those per-element nu);
 *   m_vals [nnz]  sigma-mass, sigma [ne] device or NULL (sigma = 1);
 *   j_vals [nnz]  Jacobian, as mpk_csc_jacobian;
 *   residual [ndofs] R = K(nu) a, as mpk_csc_jacobian.
 * Any output may be NULL.  FP64 / FP32_LOCAL only (MPK_ERR_ARGS for
 * FP64_EXACT); P may be NULL after mpk_plan_set_coords. */
int mpk_csc_newton_set(const mpk_plan* plan, const double* P, const int32_t* tags,
                       const double* sol_a, const mpk_material* mat, const double* sigma,
                       const mpk_quadrature* quad_k, const mpk_quadrature* quad_m,
                       int precision, double* k_vals, double* m_vals, double* j_vals,
                       double* residual, void* stream);

/* Bytes of the library's static device scratch (one copy per device, held
 * until the process ends) that mpk_csc_newton_set_noalloc works in. */
#define MPK_NOALLOC_SCRATCH_BYTES (128ull << 20)

/* mpk_csc_newton_set without any device allocation: no counterpart in the
 * reference.  For callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same
 * arguments, outputs and values as mpk_csc_newton_set; the element state and
 * tile classes live in the static scratch of MPK_NOALLOC_SCRATCH_BYTES
 * instead of the plan's lazily allocated buffers, and the tile passes run one
 * after the other on `stream`.  Limit: 48 B per element (FP64) or 24 B
 * (FP32_LOCAL) + 8 B per 64 columns must fit the scratch (about 2.7M / 5.5M
 * elements); larger plans get MPK_ERR_LIMIT.  The current device must be the
 * plan's.  Every call on a device shares that device's scratch: calls that
 * overlap on one device (other streams, other threads) race. */
int mpk_csc_newton_set_noalloc(const mpk_plan* plan, const double* P, const int32_t* tags,
                               const double* sol_a, const mpk_material* mat,
                               const double* sigma, const mpk_quadrature* quad_k,
                               const mpk_quadrature* quad_m, int precision, double* k_vals,
                               double* m_vals, double* j_vals, double* residual, void* stream);

/* Source vector F [ndofs] = sparse(I(valid), 1, V(valid), N, 1) of
 * assemble_source_kernel_mex (src/Assembler.m:78-93), summed per DOF in
 * triplet order (bit-identical); one lane per DOF, no atomics.  Js [3 x ne]
 * device; quad: Qw + ValRef. */
int mpk_csc_source(const mpk_plan* plan, const double* P, const double* Js,
                   const mpk_quadrature* quad, double* F, void* stream);

/* Emulates sparse()'s value-dependent rule on an assembled structural CSC:
 * copies the entries that are not exactly zero into (colptr_out [n+1] int64,
 * rowind_out, vals_out) (capacity nnz) and returns the new nnz.  `complex`
 * != 0 treats vals as interleaved complex128 (an entry is dropped when both
 * parts are zero).  Synchronises. */
int mpk_csc_drop_zeros(const mpk_plan* plan, const double* vals, int complex_vals,
                       int64_t* colptr_out, int32_t* rowind_out, double* vals_out,
                       int64_t* nnz_out, void* stream);

/* ========================================================================== */
/* 5. Linear algebra used by the transient Newton driver                      */
/* ========================================================================== */
/* No reference counterpart in the assembly path: the reference solves with
 * MUMPS / backslash (src/solver/LinearSolver.m:22-46), whose binaries are
 * absent.  These operate on a structural CSC pattern of a symmetric matrix
 * (device arrays). */

/* y = A x for symmetric A stored as CSC (== CSR of A). */
int mpk_spmv_sym(const int32_t* colptr, const int32_t* rowind, const double* vals, int64_t n,
                 const double* x, double* y, void* stream);

/* out = c0*v0 + c1*v1 + c2*v2 + c3*v3 over nnz values of one pattern; v1..v3 may be NULL. */
int mpk_csc_combine(int64_t nnz, double c0, const double* v0, double c1, const double* v1,
                    double c2, const double* v2, double c3, const double* v3, double* out,
                    void* stream);

/* Doubles of device workspace mpk_pcg_solve needs for n unknowns. */
int mpk_pcg_work_size(int64_t n, int64_t* size_out);

/* Jacobi-PCG on the Dirichlet-applied operator of BoundaryCondition.applyDirichlet
 * (src/BoundaryCondition.m:81-112): fixed [n] (uint8, NULL = none) removes the
 * rows and columns of fixed DOFs and puts a unit diagonal back; b's fixed
 * entries are treated as 0.  Starts from x = 0; stops when |r|/|b| <= rel_tol
 * (checked every check_every iterations) or after max_iter.  Synchronises.
 * The iterations run on an internal stream of the calling thread, ordered
 * after `stream` and joined back into it; each batch of check_every
 * iterations is captured once into a HIP graph and replayed
 * (MPK_SOLVER_GRAPHS=0: plain launches, same results).  The same holds for
 * mpk_pcg_aux_solve and mpk_cocg_aux_solve. */
int mpk_pcg_solve(const int32_t* colptr, const int32_t* rowind, const double* vals, int64_t n,
                  const uint8_t* fixed, const double* b, double* x, double* work,
                  double rel_tol, int max_iter, int check_every, int* iters_out,
                  double* rel_res_out, void* stream);

/* ========================================================================== */
/* 6. Batched assembly: many meshes x many frequencies in one launch           */
/* ========================================================================== */
/* No single reference call: the reference assembles K and M per mesh (the MEX
 * kernels) and forms K + 1j*w*M per frequency (src/solver/FrequencySolver.m:
 * 127-131); its HBFEM kernel (src/mex/assemble_hbfem_kernel_mex.cpp:25-260)
 * stacks harmonic blocks of one mesh.  Here every item is one mesh with its
 * plan (ring order, coordinates loaded with mpk_plan_set_coords), nu and
 * sigma [ne] and an output of nfreq complex128 value sets [nfreq][2*nnz]
 * (interleaved re, im), value set f = K + 1j*omegas[f]*M_sigma. */
typedef struct mpk_batch_item {
    const mpk_plan* plan;
    const double* nu;
    const double* sigma;
    double* vals_c128;
} mpk_batch_item;
typedef struct mpk_batch mpk_batch;

/* Builds the device descriptors and the tile table (copies; synchronises).
 * precision: MPK_PREC_FP64 or MPK_PREC_FP32_LOCAL.  nfreq <= 64. */
int mpk_batch_create(const mpk_batch_item* items, int64_t count, const double* omegas,
                     int32_t nfreq, const mpk_quadrature* quad_k, const mpk_quadrature* quad_m,
                     int precision, void* stream, mpk_batch** batch_out);
/* Schedule: ring gather batch (2, 3, 4 or 6) and blocks per CU of a persistent grid
 * (0: one block per 64-column tile).  Results do not depend on it. */
int mpk_batch_schedule(mpk_batch* batch, int32_t ring_batch, int32_t blocks_per_cu);
int mpk_batch_info(const mpk_batch* batch, int64_t* ntiles, int64_t* nnz_total,
                   int64_t* lds_bytes);
/* One launch over every item and frequency; no allocation, copy or
 * synchronisation (capturable in a HIP graph). */
int mpk_batch_run(const mpk_batch* batch, void* stream);
int mpk_batch_destroy(mpk_batch* batch);

/* Auxiliary (gradient) space of an edge-element system, for mpk_pcg_aux_solve.
 * Free (non-Dirichlet) nodes are numbered 0..num_nodes-1.  Device arrays. */
typedef struct mpk_aux_space {
    int64_t num_nodes;
    const int32_t* edge_nodes; /* [2n] per DOF: free-node numbers (a, b) of its edge's lower
                                  and higher end (-1: a Dirichlet node) */
    const int32_t* node_ptr;   /* [num_nodes+1] */
    const int32_t* node_refs;  /* per node its edges: dof << 1 | (node is the higher end) */
    const int32_t* ln_ptr;     /* Ln = G^T A G over the free nodes, CSR (symmetric) */
    const int32_t* ln_idx;
    const double* ln_vals;
    int32_t sweeps;            /* damped Jacobi sweeps on Ln per application (>= 1) */
    double omega;              /* their damping, e.g. 2/3 */
} mpk_aux_space;

/* Doubles of device workspace mpk_pcg_aux_solve needs. */
int mpk_pcg_aux_work_size(int64_t n, int64_t num_nodes, int64_t* size_out);

/* mpk_pcg_solve with the preconditioner M^-1 = D^-1 + G p(Ln) G^T (Jacobi plus
 * a gradient-space correction, the auxiliary-space idea of Hiptmair and Xu):
 * the near-null space of a weakly regularised curl-curl operator (gradients
 * in non-conducting regions) is what stalls plain Jacobi-PCG.  G is the
 * edge-node incidence (edge (a < b): phi_b - phi_a) restricted to free nodes,
 * Ln = G^T A G (A's curl part vanishes on gradients, so Ln is the nodal
 * matrix of A's mass part and does not change with the Newton iterate).
 * p(Ln) = aux->sweeps damped Jacobi sweeps from zero (a fixed SPD polynomial:
 * CG stays valid).  Same stopping rule, determinism and synchronisation as
 * mpk_pcg_solve. */
int mpk_pcg_aux_solve(const int32_t* colptr, const int32_t* rowind, const double* vals,
                      int64_t n, const uint8_t* fixed, const mpk_aux_space* aux, const double* b,
                      double* x, double* work, double rel_tol, int max_iter, int check_every,
                      int* iters_out, double* rel_res_out, void* stream);

/* Doubles of device workspace mpk_cocg_aux_solve needs. */
int mpk_cocg_aux_work_size(int64_t n, int64_t num_nodes, int64_t* size_out);

/* Complex symmetric solve for the time-harmonic operator the reference hands
 * to MUMPS in symmetric mode (src/solver/FrequencySolver.m:127-139,
 * LinearSolver with MumpsSymmetry = 2): A = R + 1j*I, R and I real symmetric
 * on one structural CSC pattern (vals_re, vals_im), e.g. R = K + eps*M and
 * I = omega*M_sigma.  Dirichlet DOFs as in mpk_pcg_solve.  COCG (CG with the
 * unconjugated form x^T y) preconditioned by the real operator of
 * mpk_pcg_aux_solve built from P = R + I: Jacobi on diag(R + I) plus the
 * gradient-space correction (aux: Ln = G^T (mass part of P) G; NULL = Jacobi
 * only), applied to the real and imaginary parts.  b and x are complex128
 * [n] (interleaved re, im; 16-byte aligned), work 16-byte aligned.  x starts
 * at 0; stops when |r|_2/|b|_2 <= rel_tol (checked every check_every
 * iterations) or after max_iter.  Deterministic; synchronises. */
int mpk_cocg_aux_solve(const int32_t* colptr, const int32_t* rowind, const double* vals_re,
                       const double* vals_im, int64_t n, const uint8_t* fixed,
                       const mpk_aux_space* aux, const double* b, double* x, double* work,
                       double rel_tol, int max_iter, int check_every, int* iters_out,
                       double* rel_res_out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MPK_ASSEMBLY_H */
