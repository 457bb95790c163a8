/*
 * mpk_oracle.c -- CPU restatement (FP64) of MagPackage's element-assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine in
 * magpackage_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Every function restates one piece of the reference with the same data
 * layout (MATLAB column-major, 1-based indices stored as double), the same
 * loop structure and the same floating-point operation order, so that a
 * bit-exact comparison is meaningful.  Build flags: -O2 -ffp-contract=off
 * (no FMA contraction, as the reference's x86-64 -O3 build without -march).
 *
 * Functions and the reference code they follow (paths relative to the
 * reference root):
 *   orc_generate_edges   src/Mesh.m:114-138              (Mesh.generateEdges)
 *   orc_distribute_dofs  src/DofHandler.m:72-105         (Nedelec branch)
 *   orc_jacobian_3d      src/mex/MexElemUtils.hpp:35-49  (compute_jacobian_3d)
 *   orc_inverse_3x3      src/mex/MexElemUtils.hpp:52-66  (compute_inverse_3x3)
 *   orc_curl_curl_kernel src/mex/assemble_curl_curl_kernel_mex.cpp:18-121
 *   orc_mass_kernel      src/mex/assemble_mass_kernel_mex.cpp:16-127
 *   orc_eval_nu          src/mex/MexMaterialUtils.hpp:24-87 (evaluate_nu_derivative)
 *   orc_jacobian_kernel  src/mex/assemble_jacobian_kernel_mex.cpp:7-153
 *   orc_source_kernel    src/mex/assemble_source_kernel_mex.cpp:16-120
 *   orc_scalar_laplacian_kernel  src/mex/assemble_scalar_laplacian_kernel_mex.cpp:13-105
 *   orc_coupling_kernel  src/mex/assemble_coupling_kernel_mex.cpp:13-131
 *   orc_winding_kernel   src/mex/assemble_winding_kernel_mex.cpp:13-126
 *   orc_hbfem_kernel     src/mex/assemble_hbfem_kernel_mex.cpp:25-260 (the spline
 *                        coefficient offsets passed explicitly: the MEX leaves
 *                        its coeStart uninitialised)
 *   orc_element_b        src/post/PostProcessor.m:64-126 (computeElementB)
 *   orc_sparse           src/Assembler.m:37-38  (valid = I>0 & J>0; MATLAB sparse():
 *                        duplicates summed, CSC with sorted rows, zeros dropped;
 *                        summation order = triplet order, an assumption: MATLAB's
 *                        own order is not documented)
 *   orc_boundary_edges   src/BoundaryCondition.m:9-79    (Nedelec branch)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* a1: Mesh.generateEdges                                              */
/* ------------------------------------------------------------------ */

typedef struct {
    uint64_t key;  /* (min << 32) | max, both 1-based node ids */
    int64_t idx;   /* raw row index 6*e + k (MATLAB order: start_nodes(:) is
                      column-major over the 6x Ne matrix, i.e. k fastest) */
} orc_edge_rec;

static int orc_edge_cmp(const void* a, const void* b) {
    const orc_edge_rec* x = (const orc_edge_rec*)a;
    const orc_edge_rec* y = (const orc_edge_rec*)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

/* Local edge definitions [1 2; 1 3; 1 4; 2 3; 2 4; 3 4] (src/Mesh.m:118). */
static const int ORC_EDGE_A[6] = {0, 0, 0, 1, 1, 2};
static const int ORC_EDGE_B[6] = {1, 2, 3, 2, 3, 3};

/*
 * T: 4 x ne node ids (1-based, int32, column-major: T[4e + k]).
 * edges_out: 2 x (6 ne) capacity, unique sorted (min,max) pairs, 1-based.
 * t2e_out: 6 x ne, 1-based edge ids (T2E = reshape(ic, 6, Ne)).
 * sign_out: 6 x ne, +1 iff T(a) <= T(b) (is_same = raw(:,1)==sorted(:,1)).
 * Returns the number of unique edges.
 */
int64_t orc_generate_edges(const int32_t* T, int64_t ne, int32_t* edges_out,
                           int32_t* t2e_out, int8_t* sign_out) {
    int64_t nraw = 6 * ne;
    orc_edge_rec* rec = (orc_edge_rec*)malloc(sizeof(orc_edge_rec) * (size_t)(nraw > 0 ? nraw : 1));
    for (int64_t e = 0; e < ne; e++) {
        for (int k = 0; k < 6; k++) {
            uint32_t a = (uint32_t)T[4 * e + ORC_EDGE_A[k]];
            uint32_t b = (uint32_t)T[4 * e + ORC_EDGE_B[k]];
            uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
            rec[6 * e + k].key = ((uint64_t)lo << 32) | hi;
            rec[6 * e + k].idx = 6 * e + k;
            sign_out[6 * e + k] = (a == lo) ? (int8_t)1 : (int8_t)-1;
        }
    }
    qsort(rec, (size_t)nraw, sizeof(orc_edge_rec), orc_edge_cmp);
    int64_t nedge = 0;
    for (int64_t i = 0; i < nraw; i++) {
        if (i == 0 || rec[i].key != rec[i - 1].key) {
            edges_out[2 * nedge + 0] = (int32_t)(rec[i].key >> 32);
            edges_out[2 * nedge + 1] = (int32_t)(rec[i].key & 0xffffffffu);
            nedge++;
        }
        t2e_out[rec[i].idx] = (int32_t)nedge; /* 1-based rank */
    }
    free(rec);
    return nedge;
}

/*
 * a2: DofHandler.distributeDofs, Nedelec branch.  active_elem[e] != 0 marks
 * the active elements (all active for a full-domain space).  Active edges get
 * offset + 1.. in ascending edge order (unique(...) sorts); others map to 0.
 * cell_dofs: 6 x ne output.  Returns the number of active dofs.
 */
int64_t orc_distribute_dofs(const int32_t* t2e, int64_t ne, int64_t nedge,
                            const uint8_t* active_elem, int64_t offset,
                            int32_t* entity_map, int32_t* cell_dofs) {
    uint8_t* act = (uint8_t*)calloc((size_t)(nedge + 1), 1);
    for (int64_t e = 0; e < ne; e++) {
        if (active_elem && !active_elem[e]) continue;
        for (int k = 0; k < 6; k++) {
            int32_t ed = t2e[6 * e + k];
            if (ed > 0) act[ed] = 1;
        }
    }
    int64_t n = 0;
    for (int64_t ed = 1; ed <= nedge; ed++) {
        entity_map[ed - 1] = act[ed] ? (int32_t)(offset + (++n)) : 0;
    }
    for (int64_t i = 0; i < 6 * ne; i++) {
        int32_t ed = t2e[i];
        cell_dofs[i] = ed > 0 ? entity_map[ed - 1] : 0;
    }
    free(act);
    return n;
}

/* ------------------------------------------------------------------ */
/* a4: element geometry (MexElemUtils.hpp)                              */
/* ------------------------------------------------------------------ */

/* J[i][j] row-major in J[3*i + j]; columns = x2-x1, x3-x1, x4-x1. */
double orc_jacobian_3d(const double* p, double* J) {
    double x1 = p[0], y1 = p[1], z1 = p[2];
    double x2 = p[3], y2 = p[4], z2 = p[5];
    double x3 = p[6], y3 = p[7], z3 = p[8];
    double x4 = p[9], y4 = p[10], z4 = p[11];
    J[0] = x2 - x1; J[3] = y2 - y1; J[6] = z2 - z1;
    J[1] = x3 - x1; J[4] = y3 - y1; J[7] = z3 - z1;
    J[2] = x4 - x1; J[5] = y4 - y1; J[8] = z4 - z1;
    double det = J[0] * (J[4] * J[8] - J[5] * J[7])
               - J[1] * (J[3] * J[8] - J[5] * J[6])
               + J[2] * (J[3] * J[7] - J[4] * J[6]);
    return det;
}

void orc_inverse_3x3(const double* J, double det, double* Ji) {
    double invDet = 1.0 / det;
    Ji[0] = (J[4] * J[8] - J[5] * J[7]) * invDet;
    Ji[1] = (J[2] * J[7] - J[1] * J[8]) * invDet;
    Ji[2] = (J[1] * J[5] - J[2] * J[4]) * invDet;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) * invDet;
    Ji[4] = (J[0] * J[8] - J[2] * J[6]) * invDet;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * invDet;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) * invDet;
    Ji[7] = (J[1] * J[6] - J[0] * J[7]) * invDet;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * invDet;
}

static inline void orc_gather(const double* P, const double* T, int64_t e, double* p) {
    for (int n = 0; n < 4; n++) {
        long node = (long)T[n + e * 4] - 1;
        p[n * 3 + 0] = P[node * 3 + 0];
        p[n * 3 + 1] = P[node * 3 + 1];
        p[n * 3 + 2] = P[node * 3 + 2];
    }
}

static inline void orc_write_block(int64_t e, const double* Dofs, const double* Signs,
                                   double K[6][6], double* I, double* Jo, double* V) {
    double s[6];
    long d[6];
    for (int k = 0; k < 6; k++) s[k] = Signs[k + e * 6];
    for (int k = 0; k < 6; k++) d[k] = (long)Dofs[k + e * 6];
    long base = e * 36;
    int cnt = 0;
    for (int c = 0; c < 6; c++) {
        for (int r = 0; r < 6; r++) {
            I[base + cnt] = (double)d[r];
            Jo[base + cnt] = (double)d[c];
            V[base + cnt] = K[r][c] * s[r] * s[c];
            cnt++;
        }
    }
}

/* ------------------------------------------------------------------ */
/* a5: curl-curl stiffness                                             */
/* ------------------------------------------------------------------ */
void orc_curl_curl_kernel(const double* P, const double* T, const double* Dofs,
                          const double* Signs, const double* Nu, const double* Qw,
                          int64_t nq, const double* RefCurl, int64_t ne,
                          double* I, double* Jo, double* V) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        double invDetJ = 1.0 / detJ;
        double K[6][6];
        memset(K, 0, sizeof(K));
        double nu = Nu[e];
        for (int64_t q = 0; q < nq; q++) {
            double w = Qw[q] * absDetJ;
            double B[3][6];
            long qo = q * 18;
            for (int i = 0; i < 6; i++) {
                double rc[3] = {RefCurl[i + 0 + qo], RefCurl[i + 6 + qo], RefCurl[i + 12 + qo]};
                for (int a = 0; a < 3; a++) {
                    double t = J[3 * a + 0] * rc[0] + J[3 * a + 1] * rc[1] + J[3 * a + 2] * rc[2];
                    B[a][i] = t * invDetJ;
                }
            }
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    double dot = B[0][i] * B[0][j] + B[1][i] * B[1][j] + B[2][i] * B[2][j];
                    K[i][j] += dot * nu * w;
                }
        }
        orc_write_block(e, Dofs, Signs, K, I, Jo, V);
    }
}

/* ------------------------------------------------------------------ */
/* a6: sigma-mass                                                      */
/* ------------------------------------------------------------------ */
/* coeffs may be NULL or have ncoeffs != ne: then sigma = 1 (mass_kernel_mex.cpp:34,66). */
void orc_mass_kernel(const double* P, const double* T, const double* Dofs,
                     const double* Signs, const double* Coeffs, int64_t ncoeffs,
                     const double* Qw, int64_t nq, const double* RefVal, int64_t ne,
                     double* I, double* Jo, double* V) {
    int hasC = (Coeffs != NULL) && (ncoeffs == ne);
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9], Ji[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        orc_inverse_3x3(J, detJ, Ji);
        double M[6][6];
        memset(M, 0, sizeof(M));
        double sigma = hasC ? Coeffs[e] : 1.0;
        if (fabs(sigma) > 1e-12) {
            for (int64_t q = 0; q < nq; q++) {
                double w = Qw[q] * absDetJ;
                double N[3][6];
                long qo = q * 18;
                for (int i = 0; i < 6; i++) {
                    double rn[3] = {RefVal[i + 0 + qo], RefVal[i + 6 + qo], RefVal[i + 12 + qo]};
                    for (int a = 0; a < 3; a++) {
                        /* multVecTrans: y[a] = Ji[0][a]*x0 + Ji[1][a]*x1 + Ji[2][a]*x2 */
                        N[a][i] = Ji[0 + a] * rn[0] + Ji[3 + a] * rn[1] + Ji[6 + a] * rn[2];
                    }
                }
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        double dot = N[0][i] * N[0][j] + N[1][i] * N[1][j] + N[2][i] * N[2][j];
                        M[i][j] += dot * sigma * w;
                    }
            }
        }
        orc_write_block(e, Dofs, Signs, M, I, Jo, V);
    }
}

/* ------------------------------------------------------------------ */
/* a8: B-H material evaluation                                          */
/* ------------------------------------------------------------------ */
typedef struct {
    const double* linearNu;
    const double* isNonlinear;
    const double* maxBSq;
    const double* breaks;
    const double* coefs;
    const double* splineStart;
    const double* splineCount;
    const double* splineCoeStart;
} orc_material;

void orc_eval_nu(double s, int tag, const orc_material* mat, double* nu_out, double* dnu_out) {
    int idx = tag - 1;
    double nu, dnu;
    if (mat->isNonlinear[idx] == 0.0) {
        *nu_out = mat->linearNu[idx];
        *dnu_out = 0.0;
        return;
    }
    double max_s = mat->maxBSq[idx];
    int is_extrap = (s > max_s);
    double s_eval = is_extrap ? max_s : s;
    int start = (int)mat->splineStart[idx];
    int coe_start = (int)mat->splineCoeStart[idx];
    int count = (int)mat->splineCount[idx];
    const double* b = &mat->breaks[start];
    int k = 0, lo = 0, hi = count - 1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        if (s_eval >= b[mid]) { k = mid; lo = mid + 1; }
        else { hi = mid - 1; }
    }
    if (k >= count) k = count - 1;
    if (k < 0) k = 0;
    double dx = s_eval - b[k];
    const double* c = &mat->coefs[coe_start + k * 4];
    double nu_eval = ((c[0] * dx + c[1]) * dx + c[2]) * dx + c[3];
    double dnu_eval = (3.0 * c[0] * dx + 2.0 * c[1]) * dx + c[2];
    if (is_extrap) { nu = nu_eval + dnu_eval * (s - max_s); dnu = dnu_eval; }
    else { nu = nu_eval; dnu = dnu_eval; }
    double nu0 = mat->linearNu[idx];
    if (nu > nu0) { nu = nu0; dnu = 0.0; }
    else if (nu < 1e-7) { nu = nu0; dnu = 0.0; }
    *nu_out = nu;
    *dnu_out = dnu;
}

/* Vectorised helper for tests: evaluate n points. */
void orc_eval_nu_many(const double* s, const int32_t* tags, int64_t n,
                      const double* lin, const double* isNon, const double* maxB,
                      const double* brk, const double* coe, const double* start,
                      const double* cnt, const double* coeStart, double* nu, double* dnu) {
    orc_material m = {lin, isNon, maxB, brk, coe, start, cnt, coeStart};
    for (int64_t i = 0; i < n; i++) orc_eval_nu(s[i], tags[i], &m, &nu[i], &dnu[i]);
}

/* ------------------------------------------------------------------ */
/* a7: nonlinear Jacobian + residual                                    */
/* ------------------------------------------------------------------ */
void orc_jacobian_kernel(const double* P, const double* T, const double* Dofs,
                         const double* Signs, const double* Tags, const double* SolA,
                         const double* Qw, int64_t nq, const double* RefCurl,
                         const double* lin, const double* isNon, const double* maxB,
                         const double* brk, const double* coe, const double* start,
                         const double* cnt, const double* coeStart, int calcJ, int64_t ne,
                         double* I, double* Jo, double* V, double* Ridx, double* Rval) {
    orc_material mat = {lin, isNon, maxB, brk, coe, start, cnt, coeStart};
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        double invDetJ = 1.0 / detJ;
        int dofs[6];
        double s[6], a[6];
        for (int i = 0; i < 6; i++) {
            dofs[i] = (int)Dofs[i + e * 6];
            s[i] = Signs[i + e * 6];
            int gid = dofs[i] - 1;
            a[i] = (gid >= 0) ? SolA[gid] * s[i] : 0.0;
        }
        double C[6][3], R[6] = {0, 0, 0, 0, 0, 0}, K[6][6];
        memset(K, 0, sizeof(K));
        int tag = (int)Tags[e];
        for (int64_t q = 0; q < nq; q++) {
            double w = Qw[q] * absDetJ;
            for (int i = 0; i < 6; i++) {
                double rc[3] = {RefCurl[i + 0 + q * 18], RefCurl[i + 6 + q * 18], RefCurl[i + 12 + q * 18]};
                for (int k = 0; k < 3; k++) {
                    double t = J[3 * k + 0] * rc[0] + J[3 * k + 1] * rc[1] + J[3 * k + 2] * rc[2];
                    C[i][k] = t * invDetJ;
                }
            }
            double B[3] = {0, 0, 0};
            for (int i = 0; i < 6; i++) {
                B[0] += a[i] * C[i][0];
                B[1] += a[i] * C[i][1];
                B[2] += a[i] * C[i][2];
            }
            double Bsq = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
            double nu, dnu;
            orc_eval_nu(Bsq, tag, &mat, &nu, &dnu);
            for (int i = 0; i < 6; i++) {
                double BCi = B[0] * C[i][0] + B[1] * C[i][1] + B[2] * C[i][2];
                R[i] += nu * BCi * w;
                if (calcJ) {
                    for (int j = 0; j < 6; j++) {
                        double CiCj = C[i][0] * C[j][0] + C[i][1] * C[j][1] + C[i][2] * C[j][2];
                        double BCj = B[0] * C[j][0] + B[1] * C[j][1] + B[2] * C[j][2];
                        double val = nu * CiCj + 2.0 * dnu * BCi * BCj;
                        K[i][j] += val * w;
                    }
                }
            }
        }
        for (int i = 0; i < 6; i++) {
            Ridx[e * 6 + i] = (double)dofs[i];
            Rval[e * 6 + i] = R[i] * s[i];
        }
        if (calcJ) {
            long kb = e * 36;
            int idx = 0;
            for (int j = 0; j < 6; j++)
                for (int i = 0; i < 6; i++) {
                    I[kb + idx] = (double)dofs[i];
                    Jo[kb + idx] = (double)dofs[j];
                    V[kb + idx] = K[i][j] * s[i] * s[j];
                    idx++;
                }
        }
    }
}

/* ------------------------------------------------------------------ */
/* a11: COO -> CSC with MATLAB sparse() semantics                       */
/* ------------------------------------------------------------------ */
/*
 * Triplets with I > 0 and J > 0 are kept (Assembler.m filter), ordered by
 * (column, row) with a stable two-pass counting sort, duplicates summed in
 * triplet order starting from +0.0, and entries whose sum is exactly zero
 * dropped when drop_zeros != 0.  rowind is 0-based; colptr has n + 1
 * entries.  Returns nnz (or -1 on an index outside m x n).
 */
int64_t orc_sparse(const double* I, const double* J, const double* V, int64_t nt,
                   int64_t m, int64_t n, int drop_zeros, int64_t* colptr,
                   int64_t* rowind, double* vals) {
    int64_t* cnt = (int64_t*)calloc((size_t)((m > n ? m : n) + 2), sizeof(int64_t));
    int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nt > 0 ? nt : 1));
    int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nt > 0 ? nt : 1));
    int64_t nv = 0;
    for (int64_t t = 0; t < nt; t++) {
        if (I[t] > 0 && J[t] > 0) {
            if ((int64_t)I[t] > m || (int64_t)J[t] > n) { free(cnt); free(tmp); free(ord); return -1; }
            tmp[nv++] = t;
        }
    }
    /* pass 1: stable by row */
    memset(cnt, 0, sizeof(int64_t) * (size_t)(m + 2));
    for (int64_t k = 0; k < nv; k++) cnt[(int64_t)I[tmp[k]]]++;
    for (int64_t r = 1; r <= m + 1; r++) cnt[r] += cnt[r - 1];
    for (int64_t k = 0; k < nv; k++) { int64_t r = (int64_t)I[tmp[k]] - 1; ord[cnt[r]++] = tmp[k]; }
    /* pass 2: stable by column */
    memset(cnt, 0, sizeof(int64_t) * (size_t)(n + 2));
    for (int64_t k = 0; k < nv; k++) cnt[(int64_t)J[ord[k]]]++;
    for (int64_t c = 1; c <= n + 1; c++) cnt[c] += cnt[c - 1];
    for (int64_t k = 0; k < nv; k++) { int64_t c = (int64_t)J[ord[k]] - 1; tmp[cnt[c]++] = ord[k]; }
    /* tmp now holds triplets sorted by (col,row), stable */
    int64_t nnz = 0;
    int64_t k = 0;
    for (int64_t c = 0; c < n; c++) {
        colptr[c] = nnz;
        while (k < nv && (int64_t)J[tmp[k]] - 1 == c) {
            int64_t r = (int64_t)I[tmp[k]] - 1;
            double sum = 0.0;
            while (k < nv && (int64_t)J[tmp[k]] - 1 == c && (int64_t)I[tmp[k]] - 1 == r) {
                sum += V[tmp[k]];
                k++;
            }
            if (!drop_zeros || sum != 0.0) {
                rowind[nnz] = r;
                vals[nnz] = sum;
                nnz++;
            }
        }
    }
    colptr[n] = nnz;
    free(cnt); free(tmp); free(ord);
    return nnz;
}

/*
 * orc_sparse on OpenMP threads, for timing the host path at the core count
 * of the machine: same contract and bit-identical output.  Triplets are
 * counted per column and scattered into column buckets with atomics (in any
 * order); each bucket is then sorted by (row, triplet index), which restores
 * the serial order within every (column, row) group, so duplicates are
 * summed in triplet order from +0.0 exactly as orc_sparse sums them.
 */
typedef struct { int64_t r, t; } orc_rt;
static int orc_rt_cmp(const void* x, const void* y) {
    const orc_rt* a = (const orc_rt*)x;
    const orc_rt* b = (const orc_rt*)y;
    if (a->r != b->r) return a->r < b->r ? -1 : 1;
    return a->t < b->t ? -1 : (a->t > b->t ? 1 : 0);
}
static void orc_rt_sort(orc_rt* q, int64_t k) {
    if (k > 48) {
        qsort(q, (size_t)k, sizeof(orc_rt), orc_rt_cmp);
        return;
    }
    for (int64_t i = 1; i < k; i++) {
        const orc_rt v = q[i];
        int64_t j = i - 1;
        while (j >= 0 && orc_rt_cmp(&q[j], &v) > 0) { q[j + 1] = q[j]; j--; }
        q[j + 1] = v;
    }
}

int64_t orc_sparse_par(const double* I, const double* J, const double* V, int64_t nt,
                       int64_t m, int64_t n, int drop_zeros, int64_t* colptr,
                       int64_t* rowind, double* vals) {
    int64_t* start = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    int64_t* ucnt = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    if (!start || !cur || !ucnt) { free(start); free(cur); free(ucnt); return -2; }
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int64_t t = 0; t < nt; t++) {
        if (I[t] > 0 && J[t] > 0) {
            const int64_t r = (int64_t)I[t], c = (int64_t)J[t];
            if (r > m || c > n) { bad = 1; continue; }
#pragma omp atomic
            start[c]++;
        }
    }
    if (bad) { free(start); free(cur); free(ucnt); return -1; }
    for (int64_t c = 0; c < n; c++) start[c + 1] += start[c];
    const int64_t nv = start[n];
    orc_rt* b = (orc_rt*)malloc(sizeof(orc_rt) * (size_t)(nv > 0 ? nv : 1));
    double* sv = (double*)malloc(sizeof(double) * (size_t)(nv > 0 ? nv : 1));
    if (!b || !sv) { free(start); free(cur); free(ucnt); free(b); free(sv); return -2; }
    memcpy(cur, start, sizeof(int64_t) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; t++) {
        if (I[t] > 0 && J[t] > 0) {
            const int64_t c = (int64_t)J[t] - 1;
            int64_t p;
#pragma omp atomic capture
            p = cur[c]++;
            b[p].r = (int64_t)I[t] - 1;
            b[p].t = t;
        }
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t c = 0; c < n; c++) {
        orc_rt* q = b + start[c];
        const int64_t k = start[c + 1] - start[c];
        orc_rt_sort(q, k);
        int64_t u = 0;
        for (int64_t i = 0; i < k;) {
            const int64_t r = q[i].r;
            double sum = 0.0;
            while (i < k && q[i].r == r) { sum += V[q[i].t]; i++; }
            if (!drop_zeros || sum != 0.0) {
                q[u].r = r;  /* u <= the consumed index: overwrites a used slot */
                sv[start[c] + u] = sum;
                u++;
            }
        }
        ucnt[c] = u;
    }
    colptr[0] = 0;
    for (int64_t c = 0; c < n; c++) colptr[c + 1] = colptr[c] + ucnt[c];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; c++) {
        const orc_rt* q = b + start[c];
        for (int64_t i = 0; i < ucnt[c]; i++) {
            rowind[colptr[c] + i] = q[i].r;
            vals[colptr[c] + i] = sv[start[c] + i];
        }
    }
    const int64_t nnz = colptr[n];
    free(start); free(cur); free(ucnt); free(b); free(sv);
    return nnz;
}

/*
 * K + 1j*w*M for two CSC matrices of the same n columns with sorted rows (the
 * sum FrequencySolver.m forms after assembly): the union of the two patterns,
 * re = K (0 where K has no entry), im = w * M (0 where M has none).  Two
 * passes over the columns on OpenMP threads; returns the nnz of the sum.
 */
int64_t orc_csc_add_jw(int64_t n, const int64_t* kp, const int64_t* kr, const double* kv,
                       const int64_t* mp, const int64_t* mr, const double* mv, double w,
                       int64_t* colptr, int64_t* rowind, double* re, double* im) {
    int64_t* ucnt = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    if (!ucnt) return -2;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; c++) {
        int64_t a = kp[c], b = mp[c], u = 0;
        while (a < kp[c + 1] || b < mp[c + 1]) {
            if (b >= mp[c + 1] || (a < kp[c + 1] && kr[a] < mr[b])) a++;
            else if (a >= kp[c + 1] || mr[b] < kr[a]) b++;
            else { a++; b++; }
            u++;
        }
        ucnt[c] = u;
    }
    colptr[0] = 0;
    for (int64_t c = 0; c < n; c++) colptr[c + 1] = colptr[c] + ucnt[c];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; c++) {
        int64_t a = kp[c], b = mp[c], o = colptr[c];
        while (a < kp[c + 1] || b < mp[c + 1]) {
            if (b >= mp[c + 1] || (a < kp[c + 1] && kr[a] < mr[b])) {
                rowind[o] = kr[a]; re[o] = kv[a]; im[o] = 0.0; a++;
            } else if (a >= kp[c + 1] || mr[b] < kr[a]) {
                rowind[o] = mr[b]; re[o] = 0.0; im[o] = w * mv[b]; b++;
            } else {
                rowind[o] = kr[a]; re[o] = kv[a]; im[o] = w * mv[b]; a++; b++;
            }
            o++;
        }
    }
    const int64_t nnz = colptr[n];
    free(ucnt);
    return nnz;
}

/* ------------------------------------------------------------------ */
/* BoundaryCondition.findOuterBoundaryDofs, Nedelec branch              */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t a, b, c; } orc_face;

static int orc_face_cmp(const void* x, const void* y) {
    const orc_face* f = (const orc_face*)x;
    const orc_face* g = (const orc_face*)y;
    if (f->a != g->a) return f->a < g->a ? -1 : 1;
    if (f->b != g->b) return f->b < g->b ? -1 : 1;
    if (f->c != g->c) return f->c < g->c ? -1 : 1;
    return 0;
}

static void orc_sort3(uint32_t* v) {
    uint32_t t;
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
    if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
    if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

/*
 * Faces seen once are boundary faces; their three edges (min,max) are
 * boundary edges.  is_bnd_edge[edge-1] = 1 for those.  edges: 2 x nedge
 * sorted unique pairs from orc_generate_edges.
 */
void orc_boundary_edges(const int32_t* T, int64_t ne, const int32_t* edges, int64_t nedge,
                        uint8_t* is_bnd_edge) {
    static const int F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    orc_face* f = (orc_face*)malloc(sizeof(orc_face) * (size_t)(4 * ne > 0 ? 4 * ne : 1));
    for (int64_t e = 0; e < ne; e++)
        for (int k = 0; k < 4; k++) {
            uint32_t v[3] = {(uint32_t)T[4 * e + F[k][0]], (uint32_t)T[4 * e + F[k][1]],
                             (uint32_t)T[4 * e + F[k][2]]};
            orc_sort3(v);
            f[4 * e + k].a = v[0]; f[4 * e + k].b = v[1]; f[4 * e + k].c = v[2];
        }
    qsort(f, (size_t)(4 * ne), sizeof(orc_face), orc_face_cmp);
    memset(is_bnd_edge, 0, (size_t)nedge);
    int64_t i = 0;
    while (i < 4 * ne) {
        int64_t j = i + 1;
        while (j < 4 * ne && orc_face_cmp(&f[i], &f[j]) == 0) j++;
        if (j - i == 1) {
            uint32_t pr[3][2] = {{f[i].a, f[i].b}, {f[i].a, f[i].c}, {f[i].b, f[i].c}};
            for (int k = 0; k < 3; k++) {
                /* binary search in the sorted edge list */
                int64_t lo = 0, hi = nedge - 1;
                while (lo <= hi) {
                    int64_t mid = lo + (hi - lo) / 2;
                    uint32_t ea = (uint32_t)edges[2 * mid], eb = (uint32_t)edges[2 * mid + 1];
                    if (ea == pr[k][0] && eb == pr[k][1]) { is_bnd_edge[mid] = 1; break; }
                    if (ea < pr[k][0] || (ea == pr[k][0] && eb < pr[k][1])) lo = mid + 1;
                    else hi = mid - 1;
                }
            }
        }
        i = j;
    }
    free(f);
}

/* ------------------------------------------------------------------ */
/* Source vector F = (J_s, N_i): assemble_source_kernel_mex            */
/* ------------------------------------------------------------------ */
/*
 * Js: [3 x ne] current density per element (physical space).  Output: 6 ne
 * pairs (I, V), slot 6e + i = (dof_i, F_i * s_i) (:104-114).  An element
 * whose |Js| components are all <= 1e-12 contributes zeros (:80).  The
 * integrand uses J_inv * Js dotted with the reference values (:68-76).
 */
void orc_source_kernel(const double* P, const double* T, const double* Dofs,
                       const double* Signs, const double* Js, const double* Qw, int64_t nq,
                       const double* RefVal, int64_t ne, double* I, double* V) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9], Ji[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        orc_inverse_3x3(J, detJ, Ji);
        double jp[3] = {Js[0 + e * 3], Js[1 + e * 3], Js[2 + e * 3]};
        double jt[3];
        for (int i = 0; i < 3; i++) jt[i] = Ji[3 * i + 0] * jp[0] + Ji[3 * i + 1] * jp[1] + Ji[3 * i + 2] * jp[2];
        double F[6] = {0, 0, 0, 0, 0, 0};
        if (fabs(jp[0]) > 1e-12 || fabs(jp[1]) > 1e-12 || fabs(jp[2]) > 1e-12) {
            for (int64_t q = 0; q < nq; q++) {
                double w = Qw[q] * absDetJ;
                long qo = q * 18;
                for (int i = 0; i < 6; i++) {
                    double rn[3] = {RefVal[i + 0 + qo], RefVal[i + 6 + qo], RefVal[i + 12 + qo]};
                    double val = jt[0] * rn[0] + jt[1] * rn[1] + jt[2] * rn[2];
                    F[i] += val * w;
                }
            }
        }
        for (int i = 0; i < 6; i++) {
            I[6 * e + i] = (double)(long)Dofs[i + e * 6];
            V[6 * e + i] = F[i] * Signs[i + e * 6];
        }
    }
}

/*
 * Element flux density at the centroid, B_e = C_phy' * (A(dofs) .* s) with
 * C_phy = (C_ref * J') / det (PostProcessor.m:103-123); C_ref is the constant
 * Nedelec P1 curl table.  Inactive DOFs (0) contribute 0.  B: [3 x ne].
 */
void orc_element_b(const double* P, const double* T, const double* Dofs, const double* Signs,
                   const double* sol, const double* CurlRef, int64_t ne, double* B) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double invDet = 1.0 / detJ;
        double b[3] = {0, 0, 0};
        for (int i = 0; i < 6; i++) {
            long d = (long)Dofs[i + e * 6];
            double a = d > 0 ? sol[d - 1] * Signs[i + e * 6] : 0.0;
            /* row i of C_ref * J': sum_k C_ref(i,k) J(c,k), then / det */
            for (int c = 0; c < 3; c++) {
                double cp = (CurlRef[i + 0] * J[3 * c + 0] + CurlRef[i + 6] * J[3 * c + 1] +
                             CurlRef[i + 12] * J[3 * c + 2]) * invDet;
                b[c] += cp * a;
            }
        }
        B[3 * e + 0] = b[0];
        B[3 * e + 1] = b[1];
        B[3 * e + 2] = b[2];
    }
}

/* ------------------------------------------------------------------ */
/* Nodal kernels of the A-V formulation and the winding source         */
/* ------------------------------------------------------------------ */
/* J^{-T} g for a reference row g (MexElemUtils multVecTrans). */
static inline void orc_jit(const double* Ji, const double* g, double* y) {
    for (int a = 0; a < 3; a++) y[a] = Ji[0 + a] * g[0] + Ji[3 + a] * g[1] + Ji[6 + a] * g[2];
}

/*
 * K = (sigma grad phi_i, grad phi_j) over P1 nodal functions.  Dofs: 4 x ne
 * (1-based node DOFs, as doubles); GradRef [4 x 3 x nq] at i + 4c + 12q.
 * Output slot 16 e + 4 c + r = (dofs[r], dofs[c], K[r][c]); elements with
 * |sigma| <= 1e-12 give a zero block (their indices are still written).
 */
void orc_scalar_laplacian_kernel(const double* P, const double* T, const double* Dofs,
                                 const double* Coeffs, const double* Qw, int64_t nq,
                                 const double* GradRef, int64_t ne, double* I, double* Jo,
                                 double* V) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9], Ji[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        orc_inverse_3x3(J, detJ, Ji);
        double K[4][4];
        memset(K, 0, sizeof(K));
        double sigma = Coeffs[e];
        if (fabs(sigma) > 1e-12) {
            for (int64_t q = 0; q < nq; q++) {
                double w = Qw[q] * absDetJ;
                double G[3][4];
                for (int i = 0; i < 4; i++) {
                    double g[3] = {GradRef[i + 0 + 12 * q], GradRef[i + 4 + 12 * q],
                                   GradRef[i + 8 + 12 * q]}, y[3];
                    orc_jit(Ji, g, y);
                    for (int a = 0; a < 3; a++) G[a][i] = y[a];
                }
                for (int i = 0; i < 4; i++)
                    for (int j = 0; j < 4; j++) {
                        double dot = G[0][i] * G[0][j] + G[1][i] * G[1][j] + G[2][i] * G[2][j];
                        K[i][j] += dot * sigma * w;
                    }
            }
        }
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 4; r++) {
                long k = 16 * e + 4 * c + r;
                I[k] = (double)(long)Dofs[r + 4 * e];
                Jo[k] = (double)(long)Dofs[c + 4 * e];
                V[k] = K[r][c];
            }
    }
}

/*
 * C = (sigma N_i, grad phi_j): 6 edge rows x 4 node columns.  DofsE/Signs 6 x ne,
 * DofsN 4 x ne; NedRef [6 x 3 x nq] at i + 6c + 18q, GradRef [4 x 3 x nq].
 * Output slot 24 e + 6 c + r = (dofsE[r], dofsN[c], C[r][c] * s[r]).
 */
void orc_coupling_kernel(const double* P, const double* T, const double* DofsE,
                         const double* Signs, const double* DofsN, const double* Coeffs,
                         const double* Qw, int64_t nq, const double* NedRef,
                         const double* GradRef, int64_t ne, double* I, double* Jo, double* V) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double p[12], J[9], Ji[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        orc_inverse_3x3(J, detJ, Ji);
        double Cl[6][4];
        memset(Cl, 0, sizeof(Cl));
        double sigma = Coeffs[e];
        if (fabs(sigma) > 1e-12) {
            for (int64_t q = 0; q < nq; q++) {
                double w = Qw[q] * absDetJ;
                double N[3][6], G[3][4];
                for (int i = 0; i < 6; i++) {
                    double g[3] = {NedRef[i + 0 + 18 * q], NedRef[i + 6 + 18 * q],
                                   NedRef[i + 12 + 18 * q]}, y[3];
                    orc_jit(Ji, g, y);
                    for (int a = 0; a < 3; a++) N[a][i] = y[a];
                }
                for (int i = 0; i < 4; i++) {
                    double g[3] = {GradRef[i + 0 + 12 * q], GradRef[i + 4 + 12 * q],
                                   GradRef[i + 8 + 12 * q]}, y[3];
                    orc_jit(Ji, g, y);
                    for (int a = 0; a < 3; a++) G[a][i] = y[a];
                }
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 4; j++) {
                        double dot = N[0][i] * G[0][j] + N[1][i] * G[1][j] + N[2][i] * G[2][j];
                        Cl[i][j] += dot * sigma * w;
                    }
            }
        }
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 6; r++) {
                long k = 24 * e + 6 * c + r;
                I[k] = (double)(long)DofsE[r + 6 * e];
                Jo[k] = (double)(long)DofsN[c + 4 * e];
                V[k] = Cl[r][c] * Signs[r + 6 * e];
            }
    }
}

/*
 * Winding source F_i = (Scale J_dir, N_i) over the elements whose tag equals
 * coil_id (as ints); other elements give (0, 0) pairs.  J_dir per element
 * from DirField [3 x ne] when given, else UniDir.  Output slot 6 e + i =
 * (dofs[i], F_i s[i]).
 */
void orc_winding_kernel(const double* P, const double* T, const double* Dofs,
                        const double* Signs, const double* Tags, double coil_id,
                        const double* DirField, const double* UniDir, double scale,
                        const double* Qw, int64_t nq, const double* RefVal, int64_t ne,
                        double* I, double* V) {
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        if ((int)Tags[e] != (int)coil_id) {
            for (int i = 0; i < 6; i++) I[6 * e + i] = V[6 * e + i] = 0.0;
            continue;
        }
        double p[12], J[9], Ji[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double absDetJ = fabs(detJ);
        orc_inverse_3x3(J, detJ, Ji);
        double jp[3];
        for (int a = 0; a < 3; a++) jp[a] = (DirField ? DirField[a + 3 * e] : UniDir[a]) * scale;
        double jt[3];
        for (int i = 0; i < 3; i++) jt[i] = Ji[3 * i + 0] * jp[0] + Ji[3 * i + 1] * jp[1] + Ji[3 * i + 2] * jp[2];
        double F[6] = {0, 0, 0, 0, 0, 0};
        for (int64_t q = 0; q < nq; q++) {
            double w = Qw[q] * absDetJ;
            for (int i = 0; i < 6; i++) {
                double val = jt[0] * RefVal[i + 0 + 18 * q] + jt[1] * RefVal[i + 6 + 18 * q] +
                             jt[2] * RefVal[i + 12 + 18 * q];
                F[i] += val * w;
            }
        }
        for (int i = 0; i < 6; i++) {
            I[6 * e + i] = (double)(long)Dofs[i + 6 * e];
            V[6 * e + i] = F[i] * Signs[i + 6 * e];
        }
    }
}

/* ------------------------------------------------------------------ */
/* HBFEM harmonic-balance residual and block Jacobian                  */
/* ------------------------------------------------------------------ */
/*
 * SolR/SolI [numDofs x H] (SolI may be NULL), DmatR/DmatI [Nt x H] (DmatI may
 * be NULL), PmatR/PmatI [H x Nt] (PmatI may be NULL), column-major.  Per element:
 * A_loc = Sol(dof, :) s (a DOF of 0 reads Sol at index h*numDofs - 1 for
 * h >= 1, as the MEX does, and 0 for h = 0, where the MEX reads before the
 * array); A_time = Re(A_loc Dmat^T) summed over h in order; per quadrature
 * point and time sample B = C A_time, nu(|B|^2), H = nu B,
 * Re(:, h) += w (C . H) Pmat(h, t); Ke += w mean_t(nu) C C^T.  Outputs
 * R rows/cols/real/imag at 6 H e + 6 h + k and, with requestJ, 36 H e triplets
 * (dof + h numDofs, block Ke Scalings[h] s_r s_c) at 36 H e + 36 h + 6 c + r.
 */
void orc_hbfem_kernel(const double* P, const double* T, const double* Dofs,
                      const double* Signs, const double* Tags, const double* SolR,
                      const double* SolI, int64_t numDofs, int64_t H, const double* DmatR,
                      const double* DmatI, int64_t Nt, const double* PmatR,
                      const double* PmatI, const double* Qw, int64_t nq, const double* RefCurl,
                      const double* lin, const double* isNon, const double* maxB,
                      const double* brk, const double* coe, const double* start,
                      const double* cnt, const double* coeStart, int requestJ,
                      const double* Scalings, int64_t ne, double* I, double* Jo, double* V,
                      double* Rrows, double* Rcols, double* RvalR, double* RvalI) {
    orc_material mat = {lin, isNon, maxB, brk, coe, start, cnt, coeStart};
#pragma omp parallel for schedule(static)
    for (long e = 0; e < ne; e++) {
        double* AR = (double*)malloc(sizeof(double) * 6 * H);
        double* AI = (double*)malloc(sizeof(double) * 6 * H);
        double* At = (double*)malloc(sizeof(double) * 6 * Nt);
        double* ReR = (double*)calloc((size_t)(6 * H), sizeof(double));
        double* ReI = (double*)calloc((size_t)(6 * H), sizeof(double));
        double dofs[6], s[6], Ke[36];
        memset(Ke, 0, sizeof(Ke));
        for (int k = 0; k < 6; k++) {
            dofs[k] = Dofs[k + e * 6];
            s[k] = Signs[k + e * 6];
            int gid = (int)dofs[k] - 1;
            for (int64_t h = 0; h < H; h++) {
                int64_t idx = gid + h * numDofs;
                double vr = idx >= 0 ? SolR[idx] : 0.0;
                AR[k + h * 6] = vr * s[k];
                if (SolI) {
                    double vi = idx >= 0 ? SolI[idx] : 0.0;
                    AI[k + h * 6] = vi * s[k];
                } else {
                    AI[k + h * 6] = 0.0;
                }
            }
        }
        for (int64_t t = 0; t < Nt; t++)
            for (int k = 0; k < 6; k++) {
                double sum = 0;
                for (int64_t h = 0; h < H; h++) {
                    double term = AR[k + h * 6] * DmatR[t + h * Nt];
                    if (SolI && DmatI) term -= AI[k + h * 6] * DmatI[t + h * Nt];
                    sum += term;
                }
                At[k + t * 6] = sum;
            }
        int tag = (int)Tags[e];
        double p[12], J[9];
        orc_gather(P, T, e, p);
        double detJ = orc_jacobian_3d(p, J);
        double invDetJ = 1.0 / detJ;
        double absDetJ = fabs(detJ);
        for (int64_t q = 0; q < nq; q++) {
            double w = Qw[q] * absDetJ;
            double Cp[18];
            for (int i = 0; i < 6; i++) {
                double cx = RefCurl[i + 0 + q * 18], cy = RefCurl[i + 6 + q * 18],
                       cz = RefCurl[i + 12 + q * 18];
                Cp[i + 0] = (cx * J[0] + cy * J[1] + cz * J[2]) * invDetJ;
                Cp[i + 6] = (cx * J[3] + cy * J[4] + cz * J[5]) * invDetJ;
                Cp[i + 12] = (cx * J[6] + cy * J[7] + cz * J[8]) * invDetJ;
            }
            double sum_nu = 0;
            for (int64_t t = 0; t < Nt; t++) {
                double Bx = 0, By = 0, Bz = 0;
                for (int i = 0; i < 6; i++) {
                    double v = At[i + t * 6];
                    Bx += Cp[i + 0] * v;
                    By += Cp[i + 6] * v;
                    Bz += Cp[i + 12] * v;
                }
                double bsq = Bx * Bx + By * By + Bz * Bz;
                double nu, dnu;
                orc_eval_nu(bsq, tag, &mat, &nu, &dnu);
                sum_nu += nu;
                double Hx = nu * Bx, Hy = nu * By, Hz = nu * Bz;
                for (int64_t h = 0; h < H; h++) {
                    double pr = PmatR[h + t * H];
                    for (int i = 0; i < 6; i++) {
                        double dot = Cp[i + 0] * Hx + Cp[i + 6] * Hy + Cp[i + 12] * Hz;
                        ReR[i + h * 6] += w * dot * pr;
                        if (PmatI) ReI[i + h * 6] += w * dot * PmatI[h + t * H];
                    }
                }
            }
            if (requestJ) {
                double nu_dc = sum_nu / (double)Nt;
                for (int c = 0; c < 6; c++)
                    for (int r = 0; r < 6; r++) {
                        double dot = Cp[r + 0] * Cp[c + 0] + Cp[r + 6] * Cp[c + 6] +
                                     Cp[r + 12] * Cp[c + 12];
                        Ke[r + c * 6] += w * nu_dc * dot;
                    }
            }
        }
        for (int64_t h = 0; h < H; h++)
            for (int k = 0; k < 6; k++) {
                long o = 6 * H * e + 6 * h + k;
                Rrows[o] = dofs[k];
                Rcols[o] = (double)(h + 1);
                RvalR[o] = ReR[k + h * 6] * s[k];
                RvalI[o] = ReI[k + h * 6] * s[k];
            }
        if (requestJ) {
            for (int64_t h = 0; h < H; h++) {
                double f = Scalings[h];
                for (int c = 0; c < 6; c++)
                    for (int r = 0; r < 6; r++) {
                        long o = 36 * H * e + 36 * h + 6 * c + r;
                        double val = Ke[r + c * 6] * f;
                        I[o] = dofs[r] + (double)(h * numDofs);
                        Jo[o] = dofs[c] + (double)(h * numDofs);
                        V[o] = val * s[r] * s[c];
                    }
            }
        }
        free(AR); free(AI); free(At); free(ReR); free(ReI);
    }
}
