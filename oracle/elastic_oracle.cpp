// elastic_oracle.cpp -- FP64 CPU restatement of Alamo's nodal elastic
// multigrid path.  TEST INFRASTRUCTURE: the checker for the HIP library, and
// the timed CPU baseline.  Never linked into the product.
//
// Restated from (paths relative to the reference root):
//   src/Operator/Elastic.cpp:139-379,746-910   Fapply, Diagonal, C averaging
//   src/Operator/Operator.cpp:88-154,262-434,635-657  Jacobi, transfers, residuals
//   src/Numeric/Stencil.H:41-301,588-657,900-963      finite differences
//   src/Set/Matrix4_MajorMinor.H:327-375,676-767      21-slot stiffness
//   src/Set/Matrix4_Isotropic.H:137-173, Matrix4_Diagonal.H:119-143
//   src/BC/Operator/Elastic/Constant.H:197-425         boundary rows
//   src/Solver/Nonlocal/Newton.H:144-231               Newton right-hand side
//   src/Solver/Nonlocal/Linear.H:255-290 + AMReX MLMG/MLCGSolver (restated
//   from the published AMReX algorithm; see ora_mlmg_solve).
#include "elastic_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Grid helpers
// ---------------------------------------------------------------------------
// Periodic axes of every grid made after ora_set_periodic (the geometry's
// is_periodic): the nodes 0 and n-1 of such an axis are one point, and an
// index past either face wraps to its periodic image (the ghost nodes AMReX
// fills through FillBoundary with the geometry's periodicity).
int g_per[3] = {0, 0, 0};

struct Grid {
    int n[3];
    double dx[3];
    long N;
    int per[3];
    Grid(const int* nn, const double* d) {
        for (int a = 0; a < 3; ++a) { n[a] = nn[a]; dx[a] = d ? d[a] : 1.0; per[a] = g_per[a]; }
        N = (long)n[0] * n[1] * n[2];
    }
    int wrap(int c, int a) const {
        if (!per[a]) return c;
        if (c < 0) return c + (n[a] - 1);
        if (c > n[a] - 1) return c - (n[a] - 1);
        return c;
    }
    long id(int i, int j, int k) const {
        return ((long)wrap(k, 2) * n[1] + wrap(j, 1)) * n[0] + wrap(i, 0);
    }
    // GetStencil on growPeriodicDomain: central on a periodic axis
    int sten(int a, int c) const;
};

// Stencil type per axis: GetStencil (src/Numeric/Stencil.H:41-60).
enum { ST_LO = 0, ST_HI = 1, ST_CENTRAL = 2 };

inline int sten_of(int i, int nn) { return i == 0 ? ST_HI : (i == nn - 1 ? ST_LO : ST_CENTRAL); }
int Grid::sten(int a, int c) const { return per[a] ? ST_CENTRAL : sten_of(c, n[a]); }

// First derivative of component m along axis a (src/Numeric/Stencil.H:70-191).
inline double d1(const double* f, const Grid& g, int i, int j, int k, int a, int st) {
    const int o[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int di = o[a][0], dj = o[a][1], dk = o[a][2];
    if (st == ST_LO) return (f[g.id(i, j, k)] - f[g.id(i - di, j - dj, k - dk)]) / g.dx[a];
    if (st == ST_HI) return (f[g.id(i + di, j + dj, k + dk)] - f[g.id(i, j, k)]) / g.dx[a];
    return (f[g.id(i + di, j + dj, k + dk)] - f[g.id(i - di, j - dj, k - dk)]) * 0.5 / g.dx[a];
}

// Boundary region of a node in the order of BC::Operator::Elastic::Elastic::Face
// (src/BC/Operator/Elastic/Elastic.H:35-43), classified in the order of
// Constant::operator() (src/BC/Operator/Elastic/Constant.H:373-400).
// Returns -1 for interior nodes.  nrm receives the unit normal, using the
// truncated constants of src/BC/Operator/Elastic/Elastic.H:79-80.
const double S3 = 0.57735026919, S2 = 0.70710678118;

int region_of(const Grid& g, int i, int j, int k, double nrm[3]) {
    // the faces of a periodic axis are inside growPeriodicDomain(1)
    const bool xl = !g.per[0] && i == 0, xh = !g.per[0] && i == g.n[0] - 1, yl = !g.per[1] && j == 0,
               yh = !g.per[1] && j == g.n[1] - 1, zl = !g.per[2] && k == 0, zh = !g.per[2] && k == g.n[2] - 1;
    const bool bx = xl || xh, by = yl || yh, bz = zl || zh;
    if (!(bx || by || bz)) return -1;
    const double sx = xl ? -1.0 : 1.0, sy = yl ? -1.0 : 1.0, sz = zl ? -1.0 : 1.0;
    if (bx && by && bz) {
        nrm[0] = sx * S3; nrm[1] = sy * S3; nrm[2] = sz * S3;
        return 18 + (xh ? 4 : 0) + (yh ? 2 : 0) + (zh ? 1 : 0);
    }
    if (by && bz) { nrm[0] = 0; nrm[1] = sy * S2; nrm[2] = sz * S2; return 6 + (yh ? 2 : 0) + (zh ? 1 : 0); }
    if (bz && bx) { nrm[0] = sx * S2; nrm[1] = 0; nrm[2] = sz * S2; return 10 + (zh ? 2 : 0) + (xh ? 1 : 0); }
    if (bx && by) { nrm[0] = sx * S2; nrm[1] = sy * S2; nrm[2] = 0; return 14 + (xh ? 2 : 0) + (yh ? 1 : 0); }
    nrm[0] = nrm[1] = nrm[2] = 0.0;
    if (bx) { nrm[0] = sx; return xl ? 0 : 3; }
    if (by) { nrm[1] = sy; return yl ? 1 : 4; }
    nrm[2] = sz;
    return zl ? 2 : 5;
}

// Constant::set (src/BC/Operator/Elastic/Constant.H:408-425).
// Types: 0 displacement, 1 traction, 2 periodic, 3 neumann, 4 none.
void bc_set(const int* type3, const double u[3], const double gu[3][3], const double sg[3][3],
            const double n[3], double f[3]) {
    for (int c = 0; c < 3; ++c) {
        f[c] = 0.0;
        if (type3[c] == 0) f[c] = u[c];
        else if (type3[c] == 1) f[c] = sg[c][0] * n[0] + sg[c][1] * n[1] + sg[c][2] * n[2];
        else if (type3[c] == 3) f[c] = gu[c][0] * n[0] + gu[c][1] * n[1] + gu[c][2] * n[2];
    }
}

// ---------------------------------------------------------------------------
// Stiffness contractions
// ---------------------------------------------------------------------------
// The 21 MajorMinor slots are the upper triangle, row by row, of the 6x6
// matrix indexed by the symmetric pairs (00,01,02,11,12,22): this is the
// uid map of src/Set/Matrix4_MajorMinor.H:327-375.
inline int vpair(int a, int b) {
    static const int m[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return m[a][b];
}
inline int tri6(int P, int Q) {
    if (P > Q) std::swap(P, Q);
    return P * 6 - P * (P - 1) / 2 + (Q - P);
}

int ncomp_of(int sym) { return sym == 0 ? 2 : (sym == 1 ? 21 : (sym == 2 ? 9 : -1)); }

// sigma = C : A   (Matrix4_Isotropic.H:137-157, Matrix4_MajorMinor.H:676-691,
// Matrix4_Diagonal.H:119-132).  c[] holds the node's ncomp values.
void contract2(int sym, const double* c, const double A[3][3], double s[3][3]) {
    if (sym == 0) {
        const double lam = c[0], mu = c[1];
        s[0][0] = (lam + 2. * mu) * A[0][0] + lam * A[1][1] + lam * A[2][2];
        s[1][1] = lam * A[0][0] + (lam + 2. * mu) * A[1][1] + lam * A[2][2];
        s[2][2] = lam * A[0][0] + lam * A[1][1] + (lam + 2. * mu) * A[2][2];
        s[1][2] = s[2][1] = mu * (A[1][2] + A[2][1]);
        s[2][0] = s[0][2] = mu * (A[2][0] + A[0][2]);
        s[0][1] = s[1][0] = mu * (A[0][1] + A[1][0]);
    } else if (sym == 1) {
        const double e[6] = {A[0][0], A[0][1] + A[1][0], A[0][2] + A[2][0],
                             A[1][1], A[1][2] + A[2][1], A[2][2]};
        const int rows[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
        for (int P = 0; P < 6; ++P) {
            double acc = 0.0;
            for (int Q = 0; Q < 6; ++Q) acc += c[tri6(P, Q)] * e[Q];
            s[rows[P][0]][rows[P][1]] = acc;
            s[rows[P][1]][rows[P][0]] = acc;
        }
    } else {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) s[a][b] = c[3 * a + b] * A[a][b];
    }
}

// f_i = C_ijkl B(k,l,j) with B(k,l,j) = u_{k,lj}   (Matrix4_Isotropic.H:163-173,
// Matrix4_MajorMinor.H:709-767, Matrix4_Diagonal.H:134-143).
void contract3(int sym, const double* c, const double B[3][3][3], double f[3]) {
    if (sym == 0) {
        const double lam = c[0], mu = c[1];
        for (int i = 0; i < 3; ++i) {
            double r = 0.0;
            for (int j = 0; j < 3; ++j) r += mu * (B[i][j][j] + B[j][i][j]) + lam * B[j][j][i];
            f[i] = r;
        }
    } else if (sym == 1) {
        for (int i = 0; i < 3; ++i) {
            double r = 0.0;
            for (int j = 0; j < 3; ++j) {
                const double e[6] = {B[0][0][j], B[0][1][j] + B[1][0][j], B[0][2][j] + B[2][0][j],
                                     B[1][1][j], B[1][2][j] + B[2][1][j], B[2][2][j]};
                const int P = vpair(i, j);
                for (int Q = 0; Q < 6; ++Q) r += c[tri6(P, Q)] * e[Q];
            }
            f[i] = r;
        }
    } else {
        for (int i = 0; i < 3; ++i) {
            double r = 0.0;
            for (int j = 0; j < 3; ++j) r += c[3 * i + j] * B[i][j][j];
            f[i] = r;
        }
    }
}

inline void gather_c(int sym, const double* C, long N, long id, double* c) {
    const int nc = ncomp_of(sym);
    for (int s = 0; s < nc; ++s) c[s] = C[s * N + id];
}

// ---------------------------------------------------------------------------
// Operator pieces
// ---------------------------------------------------------------------------
struct Op {
    Grid g;
    int sym;
    const double* C;
    const int* bct;  // 26 x 3
    int uniform;
};

// Elastic<SYM>::Fapply body for one node (src/Operator/Elastic.cpp:165-290).
void fapply_node(const Op& op, const double* u, int i, int j, int k, double out[3]) {
    const Grid& g = op.g;
    const long N = g.N, id = g.id(i, j, k);
    const int st[3] = {g.sten(0, i), g.sten(1, j), g.sten(2, k)};
    double uu[3], gu[3][3], c[21];
    for (int p = 0; p < 3; ++p) {
        uu[p] = u[p * N + id];
        for (int a = 0; a < 3; ++a) gu[p][a] = d1(u + p * N, g, i, j, k, a, st[a]);
    }
    gather_c(op.sym, op.C, N, id, c);
    double nrm[3];
    const int reg = region_of(g, i, j, k, nrm);
    if (reg >= 0) {
        double sg[3][3];
        contract2(op.sym, c, gu, sg);
        bc_set(op.bct + 3 * reg, uu, gu, sg, nrm, out);
        return;
    }
    // gradgradu[p](l,m) = u_{p,lm}: central 2nd derivatives and 4-point mixed
    // derivatives (src/Numeric/Stencil.H:197-301).
    double B[3][3][3];
    const double dx = g.dx[0], dy = g.dx[1], dz = g.dx[2];
    for (int p = 0; p < 3; ++p) {
        const double* f = u + p * N;
        const double c0 = f[id];
        B[p][0][0] = (f[g.id(i + 1, j, k)] - 2.0 * c0 + f[g.id(i - 1, j, k)]) / dx / dx;
        B[p][1][1] = (f[g.id(i, j + 1, k)] - 2.0 * c0 + f[g.id(i, j - 1, k)]) / dy / dy;
        B[p][2][2] = (f[g.id(i, j, k + 1)] - 2.0 * c0 + f[g.id(i, j, k - 1)]) / dz / dz;
        B[p][0][1] = 0.5 * 0.5 * (f[g.id(i + 1, j + 1, k)] + f[g.id(i - 1, j - 1, k)] -
                                  f[g.id(i + 1, j - 1, k)] - f[g.id(i - 1, j + 1, k)]) / (dx * dy);
        B[p][0][2] = 0.5 * 0.5 * (f[g.id(i + 1, j, k + 1)] + f[g.id(i - 1, j, k - 1)] -
                                  f[g.id(i + 1, j, k - 1)] - f[g.id(i - 1, j, k + 1)]) / (dx * dz);
        B[p][1][2] = 0.5 * 0.5 * (f[g.id(i, j + 1, k + 1)] + f[g.id(i, j - 1, k - 1)] -
                                  f[g.id(i, j + 1, k - 1)] - f[g.id(i, j - 1, k + 1)]) / (dy * dz);
        B[p][1][0] = B[p][0][1];
        B[p][2][0] = B[p][0][2];
        B[p][2][1] = B[p][1][2];
    }
    contract3(op.sym, c, B, out);
    if (!op.uniform) {
        // f += sum_a ((dC/dx_a) : grad u).col(a)   (Elastic.cpp:245-254)
        const int nc = ncomp_of(op.sym);
        const int o[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        double cg[21], sg[3][3];
        for (int a = 0; a < 3; ++a) {
            const long ip = g.id(i + o[a][0], j + o[a][1], k + o[a][2]);
            const long im = g.id(i - o[a][0], j - o[a][1], k - o[a][2]);
            for (int s = 0; s < nc; ++s) cg[s] = (op.C[s * N + ip] - op.C[s * N + im]) * 0.5 / g.dx[a];
            contract2(op.sym, cg, gu, sg);
            for (int r = 0; r < 3; ++r) out[r] += sg[r][a];
        }
    }
}

void fapply(const Op& op, const double* u, double* f) {
    const Grid& g = op.g;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                double o[3];
                fapply_node(op, u, i, j, k, o);
                const long id = g.id(i, j, k);
                for (int p = 0; p < 3; ++p) f[p * g.N + id] = o[p];
            }
}

// Elastic<SYM>::Diagonal (src/Operator/Elastic.cpp:321-374) with
// Gradient_Diagonal (src/Numeric/Stencil.H:900-963).
void diagonal(const Op& op, double* diag) {
    const Grid& g = op.g;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                const long id = g.id(i, j, k);
                double c[21], nrm[3];
                gather_c(op.sym, op.C, g.N, id, c);
                const int reg = region_of(g, i, j, k, nrm);
                const int st[3] = {g.sten(0, i), g.sten(1, j), g.sten(2, k)};
                for (int p = 0; p < 3; ++p) {
                    double val;
                    if (reg >= 0) {
                        double gu[3][3] = {{0}}, sg[3][3], u[3] = {0, 0, 0}, f[3];
                        for (int a = 0; a < 3; ++a) {
                            const bool mn = st[a] == ST_HI, mx = st[a] == ST_LO;
                            gu[p][a] = ((!mx ? 0.0 : 1.0) - (!mn ? 0.0 : 1.0)) /
                                       ((mn || mx ? 1.0 : 2.0) * g.dx[a]);
                        }
                        contract2(op.sym, c, gu, sg);
                        u[p] = 1.0;
                        bc_set(op.bct + 3 * reg, u, gu, sg, nrm, f);
                        val = f[p];
                    } else {
                        double B[3][3][3] = {{{0}}}, f[3];
                        for (int l = 0; l < 3; ++l) B[p][l][l] = -2.0 / g.dx[l] / g.dx[l];
                        contract3(op.sym, c, B, f);
                        val = 0.0 + f[p];
                    }
                    diag[p * g.N + id] = val;
                }
            }
}

// One Operator<Grid::Node>::Fsmooth call: two damped Jacobi sweeps
// (src/Operator/Operator.cpp:105-153).
void fsmooth(const Op& op, const double* diag, double omega, double* x, const double* b,
             std::vector<double>& Ax) {
    const long M = 3 * op.g.N;
    for (int ctr = 0; ctr < 2; ++ctr) {
        fapply(op, x, Ax.data());
#pragma omp parallel for schedule(static)
        for (long q = 0; q < M; ++q) {
            const double Dx = x[q] * diag[q];
            const double Rx = Ax[q] - Dx;
            x[q] = (1. - omega) * x[q] + omega * (b[q] - Rx) / diag[q];
        }
    }
}

void residual(const Op& op, const double* x, const double* b, double* r) {
    const long M = 3 * op.g.N;
    fapply(op, x, r);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < M; ++q) r[q] = b[q] + (-1.0) * r[q];
}

// Nodal full-weighting restriction (src/Operator/Operator.cpp:280-353); also
// the coefficient averaging of src/Operator/Elastic.cpp:786-834, whose
// weights are identical.
void restrict_field(const int* nf, int ncomp, const double* fine, double* crse) {
    const int nc[3] = {(nf[0] - 1) / 2 + 1, (nf[1] - 1) / 2 + 1, (nf[2] - 1) / 2 + 1};
    const long Nf = (long)nf[0] * nf[1] * nf[2], Nc = (long)nc[0] * nc[1] * nc[2];
    const int hx = nc[0] - 1, hy = nc[1] - 1, hz = nc[2] - 1;
    const Grid gf(nf, nullptr);  // the fine grid, wrapping on the periodic axes
    for (int n = 0; n < ncomp; ++n) {
        const double* fd = fine + n * Nf;
        double* cd = crse + n * Nc;
#pragma omp parallel for collapse(2) schedule(static)
        for (int K = 0; K <= hz; ++K)
            for (int J = 0; J <= hy; ++J)
                for (int I = 0; I <= hx; ++I) {
                    const int i = 2 * I, j = 2 * J, k = 2 * K;
                    auto F = [&](int a, int b, int c) { return fd[gf.id(a, b, c)]; };
                    // the coarse domain is grown along a periodic axis
                    // (src/Operator/Operator.cpp:268): full weights there
                    const bool bi = !gf.per[0] && (I == 0 || I == hx), bj = !gf.per[1] && (J == 0 || J == hy),
                               bk = !gf.per[2] && (K == 0 || K == hz);
                    double v;
                    if (bi && bj && bk) v = F(i, j, k);
                    else if (bj && bk) v = 0.25 * F(i - 1, j, k) + 0.5 * F(i, j, k) + 0.25 * F(i + 1, j, k);
                    else if (bk && bi) v = 0.25 * F(i, j - 1, k) + 0.5 * F(i, j, k) + 0.25 * F(i, j + 1, k);
                    else if (bi && bj) v = 0.25 * F(i, j, k - 1) + 0.5 * F(i, j, k) + 0.25 * F(i, j, k + 1);
                    else if (bi)
                        v = (+F(i, j - 1, k - 1) + 2.0 * F(i, j, k - 1) + F(i, j + 1, k - 1) + 2.0 * F(i, j - 1, k) +
                             4.0 * F(i, j, k) + 2.0 * F(i, j + 1, k) + F(i, j - 1, k + 1) + 2.0 * F(i, j, k + 1) +
                             F(i, j + 1, k + 1)) / 16.0;
                    else if (bj)
                        v = (+F(i - 1, j, k - 1) + 2.0 * F(i - 1, j, k) + F(i - 1, j, k + 1) + 2.0 * F(i, j, k - 1) +
                             4.0 * F(i, j, k) + 2.0 * F(i, j, k + 1) + F(i + 1, j, k - 1) + 2.0 * F(i + 1, j, k) +
                             F(i + 1, j, k + 1)) / 16.0;
                    else if (bk)
                        v = (+F(i - 1, j - 1, k) + 2.0 * F(i, j - 1, k) + F(i + 1, j - 1, k) + 2.0 * F(i - 1, j, k) +
                             4.0 * F(i, j, k) + 2.0 * F(i + 1, j, k) + F(i - 1, j + 1, k) + 2.0 * F(i, j + 1, k) +
                             F(i + 1, j + 1, k)) / 16.0;
                    else
                        v = (F(i - 1, j - 1, k - 1) + F(i - 1, j - 1, k + 1) + F(i - 1, j + 1, k - 1) +
                             F(i - 1, j + 1, k + 1) + F(i + 1, j - 1, k - 1) + F(i + 1, j - 1, k + 1) +
                             F(i + 1, j + 1, k - 1) + F(i + 1, j + 1, k + 1)) / 64.0 +
                            (F(i, j - 1, k - 1) + F(i, j - 1, k + 1) + F(i, j + 1, k - 1) + F(i, j + 1, k + 1) +
                             F(i - 1, j, k - 1) + F(i + 1, j, k - 1) + F(i - 1, j, k + 1) + F(i + 1, j, k + 1) +
                             F(i - 1, j - 1, k) + F(i - 1, j + 1, k) + F(i + 1, j - 1, k) + F(i + 1, j + 1, k)) / 32.0 +
                            (F(i - 1, j, k) + F(i, j - 1, k) + F(i, j, k - 1) + F(i + 1, j, k) + F(i, j + 1, k) +
                             F(i, j, k + 1)) / 16.0 +
                            F(i, j, k) / 8.0;
                    cd[((long)K * nc[1] + J) * nc[0] + I] = v;
                }
    }
}

// Nodal injection, crse(I,J,K) = fine(2I,2J,2K): amrex::average_down on nodal
// data (amrex_avgdown_nodes).
void inject_field(const int* nf, int ncomp, const double* fine, double* crse) {
    const int nc[3] = {(nf[0] - 1) / 2 + 1, (nf[1] - 1) / 2 + 1, (nf[2] - 1) / 2 + 1};
    const long Nf = (long)nf[0] * nf[1] * nf[2], Nc = (long)nc[0] * nc[1] * nc[2];
    for (int n = 0; n < ncomp; ++n)
#pragma omp parallel for collapse(2) schedule(static)
        for (int K = 0; K < nc[2]; ++K)
            for (int J = 0; J < nc[1]; ++J)
                for (int I = 0; I < nc[0]; ++I)
                    crse[n * Nc + ((long)K * nc[1] + J) * nc[0] + I] =
                        fine[n * Nf + ((long)(2 * K) * nf[1] + 2 * J) * nf[0] + 2 * I];
}

// Trilinear nodal prolongation, fine += P crse (src/Operator/Operator.cpp:399-428).
void interp_add(const int* nf, int ncomp, const double* crse, double* fine) {
    const int nc[3] = {(nf[0] - 1) / 2 + 1, (nf[1] - 1) / 2 + 1, (nf[2] - 1) / 2 + 1};
    const long Nf = (long)nf[0] * nf[1] * nf[2], Nc = (long)nc[0] * nc[1] * nc[2];
    for (int n = 0; n < ncomp; ++n) {
        const double* cd = crse + n * Nc;
        double* fd = fine + n * Nf;
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 0; k < nf[2]; ++k)
            for (int j = 0; j < nf[1]; ++j)
                for (int i = 0; i < nf[0]; ++i) {
                    const int I = i / 2, J = j / 2, K = k / 2;
                    auto Cc = [&](int a, int b, int c) { return cd[((long)c * nc[1] + b) * nc[0] + a]; };
                    double v;
                    if (i % 2 == 0 && j % 2 == 0 && k % 2 == 0) v = Cc(I, J, K);
                    else if (j % 2 == 0 && k % 2 == 0) v = 0.5 * (Cc(I, J, K) + Cc(I + 1, J, K));
                    else if (k % 2 == 0 && i % 2 == 0) v = 0.5 * (Cc(I, J, K) + Cc(I, J + 1, K));
                    else if (i % 2 == 0 && j % 2 == 0) v = 0.5 * (Cc(I, J, K) + Cc(I, J, K + 1));
                    else if (i % 2 == 0)
                        v = 0.25 * (Cc(I, J, K) + Cc(I, J + 1, K) + Cc(I, J, K + 1) + Cc(I, J + 1, K + 1));
                    else if (j % 2 == 0)
                        v = 0.25 * (Cc(I, J, K) + Cc(I, J, K + 1) + Cc(I + 1, J, K) + Cc(I + 1, J, K + 1));
                    else if (k % 2 == 0)
                        v = 0.25 * (Cc(I, J, K) + Cc(I + 1, J, K) + Cc(I, J + 1, K) + Cc(I + 1, J + 1, K));
                    else
                        v = 0.125 * (Cc(I, J, K) + Cc(I + 1, J, K) + Cc(I, J + 1, K) + Cc(I, J, K + 1) +
                                     Cc(I, J + 1, K + 1) + Cc(I + 1, J, K + 1) + Cc(I + 1, J + 1, K) +
                                     Cc(I + 1, J + 1, K + 1));
                    fd[((long)k * nf[1] + j) * nf[0] + i] += v;
                }
    }
}

double norm_inf(const double* v, long M) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long q = 0; q < M; ++q) m = std::max(m, std::fabs(v[q]));
    return m;
}

double dot(const double* a, const double* b, long M) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long q = 0; q < M; ++q) s += a[q] * b[q];
    return s;
}

// ---------------------------------------------------------------------------
// Multigrid hierarchy and cycles
// ---------------------------------------------------------------------------
struct MG {
    int nlev = 0;
    std::vector<Grid> g;
    std::vector<std::vector<double>> C, diag, cor, cor_hold, res, rescor, tmp;
    int sym, uniform, nu1, nu2, nuf, bottom_max_iter, cycle;
    const int* bct;
    double omega, bottom_tol_rel;

    Op op(int l) const { return Op{g[l], sym, C[l].data(), bct, uniform}; }
    long M(int l) const { return 3 * g[l].N; }

    void smooth(int l, double* x, const double* b, int n) {
        const Op o = op(l);
        for (int s = 0; s < n; ++s) fsmooth(o, diag[l].data(), omega, x, b, tmp[l]);
    }
    void corr_res(int l, double* r, const double* x, const double* b) { residual(op(l), x, b, r); }

    // MLCGSolver BiCGStab with Operator<Grid::Node>::normalize (division by
    // the diagonal, src/Operator/Operator.cpp:156-165) as left preconditioner.
    // Restated from AMReX MLCGSolver::solve_bicgstab.  Returns the AMReX code.
    int bicgstab(int l, double* sol, const double* rhs) {
        const long m = M(l);
        const Op o = op(l);
        const double* D = diag[l].data();
        std::vector<double> ph(m), sh(m), p(m), r(m), s(m), rh(m), v(m), t(m), sorig(m);
        residual(o, sol, rhs, r.data());
        for (long q = 0; q < m; ++q) r[q] /= D[q];
        sorig.assign(sol, sol + m);
        rh = r;
        std::fill(sol, sol + m, 0.0);
        double rnorm = norm_inf(r.data(), m);
        const double rnorm0 = rnorm, eps_rel = bottom_tol_rel, eps_abs = -1.0;
        int ret = 0;
        double rho_1 = 0, alpha = 0, omg = 0;
        if (rnorm0 == 0 || rnorm0 < eps_abs) return 0;
        int iter = 1;
        for (; iter <= bottom_max_iter; ++iter) {
            const double rho = dot(rh.data(), r.data(), m);
            if (rho == 0) { ret = 1; break; }
            if (iter == 1) p = r;
            else {
                const double beta = (rho / rho_1) * (alpha / omg);
                for (long q = 0; q < m; ++q) p[q] += -omg * v[q];
                for (long q = 0; q < m; ++q) p[q] = r[q] + beta * p[q];
            }
            ph = p;
            fapply(o, ph.data(), v.data());
            for (long q = 0; q < m; ++q) v[q] /= D[q];
            const double rhTv = dot(rh.data(), v.data(), m);
            if (rhTv != 0.0) alpha = rho / rhTv;
            else { ret = 2; break; }
            for (long q = 0; q < m; ++q) sol[q] += alpha * ph[q];
            for (long q = 0; q < m; ++q) s[q] = 1.0 * r[q] + (-alpha) * v[q];
            rnorm = norm_inf(s.data(), m);
            if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
            sh = s;
            fapply(o, sh.data(), t.data());
            for (long q = 0; q < m; ++q) t[q] /= D[q];
            const double tt = dot(t.data(), t.data(), m), ts = dot(t.data(), s.data(), m);
            if (tt != 0.0) omg = ts / tt;
            else { ret = 3; break; }
            for (long q = 0; q < m; ++q) sol[q] += omg * sh[q];
            for (long q = 0; q < m; ++q) r[q] = 1.0 * s[q] + (-omg) * t[q];
            rnorm = norm_inf(r.data(), m);
            if (rnorm < eps_rel * rnorm0 || rnorm < eps_abs) break;
            if (omg == 0) { ret = 4; break; }
            rho_1 = rho;
        }
        if (ret == 0 && rnorm > eps_rel * rnorm0 && rnorm > eps_abs) ret = 8;
        if ((ret == 0 || ret == 8) && (rnorm < rnorm0)) {
            for (long q = 0; q < m; ++q) sol[q] += sorig[q];
        } else {
            for (long q = 0; q < m; ++q) sol[q] = sorig[q];
        }
        return ret;
    }

    // MLMG::actualBottomSolve with the BiCGStab bottom solver (Linear.H:258).
    void bottom_solve() {
        const int l = nlev - 1;
        double* x = cor[l].data();
        std::fill(x, x + M(l), 0.0);
        const int ret = bicgstab(l, x, res[l].data());
        if (ret != 0 && ret != 9) std::fill(x, x + M(l), 0.0);
        const int n = (ret == 0) ? 0 : nuf;
        smooth(l, x, res[l].data(), n);
    }

    // MLMG::mgVcycle on AMR level 0 from MG level top.
    void vcycle(int top) {
        const int bot = nlev - 1;
        for (int l = top; l < bot; ++l) {
            std::fill(cor[l].begin(), cor[l].end(), 0.0);
            smooth(l, cor[l].data(), res[l].data(), nu1);
            corr_res(l, rescor[l].data(), cor[l].data(), res[l].data());
            restrict_field(g[l].n, 3, rescor[l].data(), res[l + 1].data());
        }
        bottom_solve();
        for (int l = bot - 1; l >= top; --l) {
            interp_add(g[l].n, 3, cor[l + 1].data(), cor[l].data());
            smooth(l, cor[l].data(), res[l].data(), nu2);
        }
    }

    // MLMG::mgFcycle.  The residual goes down by MLLinOp::avgDownResMG, i.e.
    // amrex::average_down, which injects nodal data (fcycle_inject = 1), or
    // by the operator's restriction (fcycle_inject = 0).
    int fcycle_inject = 1;
    void fcycle() {
        const int bot = nlev - 1;
        for (int l = 1; l <= bot; ++l) {
            if (fcycle_inject) inject_field(g[l - 1].n, 3, res[l - 1].data(), res[l].data());
            else restrict_field(g[l - 1].n, 3, res[l - 1].data(), res[l].data());
        }
        bottom_solve();
        for (int l = bot - 1; l >= 0; --l) {
            std::fill(cor[l].begin(), cor[l].end(), 0.0);
            interp_add(g[l].n, 3, cor[l + 1].data(), cor[l].data());
            corr_res(l, rescor[l].data(), cor[l].data(), res[l].data());
            res[l] = rescor[l];
            std::swap(cor[l], cor_hold[l]);
            vcycle(l);
            for (long q = 0; q < M(l); ++q) cor[l][q] += cor_hold[l][q];
        }
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

int ora_ncomp_of_sym(int sym) { return ncomp_of(sym); }

// The periodic axes of every later call (process-wide; 0 0 0 restores the
// default of no periodic axis).
int ora_set_periodic(const int per[3]) {
    for (int a = 0; a < 3; ++a) g_per[a] = per[a] ? 1 : 0;
    return 0;
}

int ora_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int ora_fapply(const int n[3], const double dx[3], int sym, const double* C, const int* bct, int uniform,
               const double* u, double* f) {
    if (ncomp_of(sym) < 0) return -1;
    Op op{Grid(n, dx), sym, C, bct, uniform};
    fapply(op, u, f);
    return 0;
}

int ora_diagonal(const int n[3], const double dx[3], int sym, const double* C, const int* bct, double* diag) {
    if (ncomp_of(sym) < 0) return -1;
    Op op{Grid(n, dx), sym, C, bct, 0};
    diagonal(op, diag);
    return 0;
}

int ora_fsmooth(const int n[3], const double dx[3], int sym, const double* C, const int* bct, int uniform,
                const double* diag, double omega, int nsmooth, double* x, const double* b) {
    if (ncomp_of(sym) < 0) return -1;
    Op op{Grid(n, dx), sym, C, bct, uniform};
    std::vector<double> Ax(3 * op.g.N);
    for (int s = 0; s < nsmooth; ++s) fsmooth(op, diag, omega, x, b, Ax);
    return 0;
}

int ora_residual(const int n[3], const double dx[3], int sym, const double* C, const int* bct, int uniform,
                 const double* x, const double* b, double* r) {
    if (ncomp_of(sym) < 0) return -1;
    Op op{Grid(n, dx), sym, C, bct, uniform};
    residual(op, x, b, r);
    return 0;
}

int ora_restrict(const int nf[3], int ncomp, const double* fine, double* crse) {
    for (int a = 0; a < 3; ++a)
        if (nf[a] < 3 || (nf[a] - 1) % 2) return -1;
    restrict_field(nf, ncomp, fine, crse);
    return 0;
}

int ora_interp_add(const int nf[3], int ncomp, const double* crse, double* fine) {
    for (int a = 0; a < 3; ++a)
        if (nf[a] < 3 || (nf[a] - 1) % 2) return -1;
    interp_add(nf, ncomp, crse, fine);
    return 0;
}

int ora_average_down_C(const int nf[3], int sym, const double* Cf, double* Cc) {
    if (ncomp_of(sym) < 0) return -1;
    return ora_restrict(nf, ncomp_of(sym), Cf, Cc);
}

int ora_bc_init(const int n[3], const int* bct, const double* bcv, int homogeneous, double* rhs) {
    Grid g(n, nullptr);
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                double nrm[3];
                const int reg = region_of(g, i, j, k, nrm);
                if (reg < 0) continue;
                for (int d = 0; d < 3; ++d) {
                    if (homogeneous && bct[3 * reg + d] == 0) rhs[d * g.N + g.id(i, j, k)] = 0.0;
                    else rhs[d * g.N + g.id(i, j, k)] = bcv[3 * reg + d];
                }
            }
    return 0;
}

// Newton<T>::prepareForSolve, vector version (src/Solver/Nonlocal/Newton.H:151-231).
int ora_newton_rhs(const int n[3], const double dx[3], int sym, const double* C, const double* F0,
                   const int* bct, const double* u, const double* b, double* rhs) {
    if (ncomp_of(sym) < 0) return -1;
    Grid g(n, dx);
    const long N = g.N;
    std::vector<double> dw(9 * N);
    std::vector<double> gus(9 * N);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                const long id = g.id(i, j, k);
                const int st[3] = {g.sten(0, i), g.sten(1, j), g.sten(2, k)};
                double gu[3][3], A[3][3], s[3][3], c[21];
                for (int p = 0; p < 3; ++p)
                    for (int a = 0; a < 3; ++a) gu[p][a] = d1(u + p * N, g, i, j, k, a, st[a]);
                for (int p = 0; p < 3; ++p)
                    for (int a = 0; a < 3; ++a) A[p][a] = gu[p][a] - (F0 ? F0[(3 * p + a) * N + id] : 0.0);
                gather_c(sym, C, N, id, c);
                contract2(sym, c, F0 ? A : gu, s);
                for (int p = 0; p < 3; ++p)
                    for (int a = 0; a < 3; ++a) {
                        dw[(3 * p + a) * N + id] = s[p][a];
                        gus[(3 * p + a) * N + id] = gu[p][a];
                    }
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                const long id = g.id(i, j, k);
                double nrm[3];
                const int reg = region_of(g, i, j, k, nrm);
                if (reg >= 0) {
                    double uu[3], gu[3][3], s[3][3], f[3];
                    for (int p = 0; p < 3; ++p) {
                        uu[p] = u[p * N + id];
                        for (int a = 0; a < 3; ++a) {
                            gu[p][a] = gus[(3 * p + a) * N + id];
                            s[p][a] = dw[(3 * p + a) * N + id];
                        }
                    }
                    bc_set(bct + 3 * reg, uu, gu, s, nrm, f);
                    for (int d = 0; d < 3; ++d) rhs[d * N + id] = b[d * N + id] - f[d];
                } else {
                    // Numeric::Divergence, central (src/Numeric/Stencil.H:588-657)
                    double dv[3] = {0, 0, 0};
                    const int o[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                    for (int a = 0; a < 3; ++a) {
                        const long ip = g.id(i + o[a][0], j + o[a][1], k + o[a][2]);
                        const long im = g.id(i - o[a][0], j - o[a][1], k - o[a][2]);
                        for (int r = 0; r < 3; ++r)
                            dv[r] += (dw[(3 * r + a) * N + ip] - dw[(3 * r + a) * N + im]) / 2. / g.dx[a];
                    }
                    for (int d = 0; d < 3; ++d) rhs[d * N + id] = b[d * N + id] - dv[d];
                }
            }
    return 0;
}

int ora_stress(const int n[3], const double dx[3], int sym, const double* C, const double* F0, const double* u,
               double* sigma) {
    if (ncomp_of(sym) < 0) return -1;
    Grid g(n, dx);
    const long N = g.N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < g.n[2]; ++k)
        for (int j = 0; j < g.n[1]; ++j)
            for (int i = 0; i < g.n[0]; ++i) {
                const long id = g.id(i, j, k);
                const int st[3] = {g.sten(0, i), g.sten(1, j), g.sten(2, k)};
                double A[3][3], s[3][3], c[21];
                for (int p = 0; p < 3; ++p)
                    for (int a = 0; a < 3; ++a)
                        A[p][a] = d1(u + p * N, g, i, j, k, a, st[a]) - (F0 ? F0[(3 * p + a) * N + id] : 0.0);
                gather_c(sym, C, N, id, c);
                contract2(sym, c, A, s);
                for (int p = 0; p < 3; ++p)
                    for (int a = 0; a < 3; ++a) sigma[(3 * p + a) * N + id] = s[p][a];
            }
    return 0;
}

int ora_mlmg_solve(const int n[3], const double dx[3], int sym, const double* C, const int* bct, int uniform,
                   const int* opts, const double* ropts, double* x, const double* b, int* out_iters,
                   double* out_res, double* out_hist) {
    if (ncomp_of(sym) < 0) return -1;
    MG mg;
    const int max_iter = opts[0];
    mg.nu1 = opts[1];
    mg.nu2 = opts[2];
    mg.cycle = opts[3];
    mg.bottom_max_iter = opts[4];
    mg.nuf = opts[5];
    const int max_coarsening = opts[6];
    mg.fcycle_inject = opts[7];
    // MLMG::setMaxFmgIter / setFixedIter (src/Solver/Nonlocal/Linear.H:266-267):
    // F-cycles while iter < max_fmg_iter (always when < 0, the reference's
    // 1e8), V-cycles after; fixed_iter > 0 caps the iterations and a solve
    // that has not converged by then ends without failure
    const int max_fmg_iter = opts[8], fixed_iter = opts[9];
    const double tol_rel = ropts[0], tol_abs = ropts[1];
    mg.omega = ropts[2];
    mg.bottom_tol_rel = ropts[3];
    mg.sym = sym;
    mg.uniform = uniform;
    mg.bct = bct;
    // MG hierarchy: coarsen by 2 while every axis keeps at least 2 cells
    // (AMReX MLLinOp default mg_domain_min_width = 2).
    Grid g0(n, dx);
    mg.g.push_back(g0);
    while (max_coarsening < 0 || (int)mg.g.size() <= max_coarsening) {
        const Grid& f = mg.g.back();
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const int cells = f.n[a] - 1;
            if (cells % 2 != 0 || cells / 2 < 2) ok = false;
        }
        if (!ok) break;
        int nc[3];
        double dc[3];
        for (int a = 0; a < 3; ++a) { nc[a] = (f.n[a] - 1) / 2 + 1; dc[a] = 2.0 * f.dx[a]; }
        mg.g.push_back(Grid(nc, dc));
    }
    mg.nlev = (int)mg.g.size();
    const int nc = ncomp_of(sym);
    mg.C.resize(mg.nlev);
    mg.C[0].assign(C, C + (long)nc * g0.N);
    for (int l = 1; l < mg.nlev; ++l) {
        mg.C[l].resize((long)nc * mg.g[l].N);
        restrict_field(mg.g[l - 1].n, nc, mg.C[l - 1].data(), mg.C[l].data());
    }
    for (auto* v : {&mg.diag, &mg.cor, &mg.cor_hold, &mg.res, &mg.rescor, &mg.tmp}) v->resize(mg.nlev);
    for (int l = 0; l < mg.nlev; ++l) {
        const long m = mg.M(l);
        for (auto* v : {&mg.diag, &mg.cor, &mg.cor_hold, &mg.res, &mg.rescor, &mg.tmp}) (*v)[l].assign(m, 0.0);
        diagonal(mg.op(l), mg.diag[l].data());
    }
    const long m0 = mg.M(0);
    residual(mg.op(0), x, b, mg.res[0].data());
    const double resnorm0 = norm_inf(mg.res[0].data(), m0);
    const double rhsnorm0 = norm_inf(b, m0);
    const double max_norm = std::max(resnorm0, rhsnorm0);
    const double res_target = std::max(tol_abs, std::max(tol_rel, 1.e-16) * max_norm);
    if (out_res) { out_res[1] = resnorm0; out_res[2] = rhsnorm0; out_res[0] = resnorm0; }
    if (out_hist) out_hist[0] = resnorm0;
    if (out_iters) *out_iters = 0;
    if (resnorm0 <= res_target) return 0;
    const int limit = fixed_iter > 0 ? std::min(max_iter, fixed_iter) : max_iter;
    for (int iter = 0; iter < limit; ++iter) {
        if (mg.cycle == 0 && (max_fmg_iter < 0 || iter < max_fmg_iter)) mg.fcycle();
        else mg.vcycle(0);
        for (long q = 0; q < m0; ++q) x[q] += mg.cor[0][q];
        residual(mg.op(0), x, b, mg.res[0].data());
        const double nrm = norm_inf(mg.res[0].data(), m0);
        if (out_hist) out_hist[iter + 1] = nrm;
        if (out_iters) *out_iters = iter + 1;
        if (out_res) out_res[0] = nrm;
        if (nrm <= res_target) return 0;
    }
    return fixed_iter > 0 ? 0 : 1;
}

}  // extern "C"
