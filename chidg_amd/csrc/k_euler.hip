// Spatial residual of the DG Euler system on the device: R(q) for the
// periodic uniform hex box (mod_spatial.f90:40-197 assembles it in the
// reference: volume operators, then boundary/face operators, per element).
//
// The discretization is the host generator's (euler_jac.cpp): 5 conservative
// fields, tensor Legendre modal basis of degree P (mode m = i + (P+1)(j +
// (P+1)k), block index f*nterms + m), Gauss-Legendre quadrature with P+2
// points per direction, R_e = sum_faces int phi Fhat.n dS - int grad(phi).F dV,
// Fhat = central flux + Roe dissipation with the entropy fix of
// euler_roe_operator.f90:200-215.
//
// One workgroup per element.  The element's and its six neighbours'
// coefficients are staged in LDS; threads evaluate the state, the weighted
// Euler fluxes at the nq^3 volume points and the weighted Roe fluxes at the
// 6 nq^2 face points; then each thread owns (field, mode) outputs and
// contracts them against the basis gradients (volume) and face values
// (faces), in a fixed order.  All FP64.
#include "cdg_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr double EG = 1.4;     // gamma
constexpr int ENF = 5;
constexpr int kEulerThreads = 256;

struct EulerTables {
    double L[4][5], dL[4][5];  // [mode index k][node a]: P_k and P_k' at the nodes
    double Lf[2][4];           // [side][k]: P_k at xi = -1, +1
    double w[5];               // node weights
};

struct EulerBasis {
    EulerTables t;
    int P1, nq, nx, ny, nz;
    double detJ, dxi, dS;
};

__device__ __forceinline__ void eflux(const double (&q)[ENF], int d, double (&F)[ENF]) {
    const double u[3] = {q[1] / q[0], q[2] / q[0], q[3] / q[0]};
    const double p = (EG - 1.0) * (q[4] - 0.5 * (q[1] * u[0] + q[2] * u[1] + q[3] * u[2]));
    F[0] = q[1 + d];
    for (int k = 0; k < 3; ++k) F[1 + k] = q[1 + k] * u[d];
    F[1 + d] = F[1 + d] + p;
    F[4] = (q[4] + p) * u[d];
}

// central flux + Roe dissipation across an axis-aligned face (normal nd along ax)
__device__ void eroe(const double (&qm)[ENF], const double (&qp)[ENF], int ax, double nd, double (&Fh)[ENF]) {
    double Fm[ENF], Fp[ENF];
    eflux(qm, ax, Fm);
    eflux(qp, ax, Fp);
    double nv[3] = {0, 0, 0};
    nv[ax] = nd;
    const double rm = qm[0], rp = qp[0];
    const double um[3] = {qm[1] / rm, qm[2] / rm, qm[3] / rm}, up[3] = {qp[1] / rp, qp[2] / rp, qp[3] / rp};
    const double pm = (EG - 1.0) * (qm[4] - 0.5 * (qm[1] * um[0] + qm[2] * um[1] + qm[3] * um[2]));
    const double pp = (EG - 1.0) * (qp[4] - 0.5 * (qp[1] * up[0] + qp[2] * up[1] + qp[3] * up[2]));
    const double Hm = (qm[4] + pm) / rm, Hp = (qp[4] + pp) / rp;
    const double srm = sqrt(rm), srp = sqrt(rp), ss = srm + srp;
    const double rt = sqrt(rm * rp);
    double ut[3];
    for (int k = 0; k < 3; ++k) ut[k] = (srm * um[k] + srp * up[k]) / ss;
    const double Ht = (srm * Hm + srp * Hp) / ss;
    const double vnt = nd * ut[ax];
    const double qt2 = ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2];
    const double ct = sqrt((EG - 1.0) * (Ht - 0.5 * qt2));
    const double ct2 = ct * ct;
    const double dr = rm - rp, dp = pm - pp;
    double du[3];
    for (int k = 0; k < 3; ++k) du[k] = um[k] - up[k];
    const double dvn = nd * du[ax];
    auto fix = [&](double lam) {   // entropy fix, eps = 0.01 (euler_roe_operator.f90:200-215)
        lam = fabs(lam);
        const double e = 0.01 * ct;
        if (lam < e) return 0.5 * (0.01 * ct + lam * lam / (0.01 * ct));
        return lam;
    };
    const double l1 = fix(vnt - ct), l2 = fix(vnt), l3 = fix(vnt + ct);
    const double C1 = l1 * (dp - rt * ct * dvn) / (2.0 * ct2);
    const double C2a = l2 * (dr - dp / ct2);
    const double C2b = l2 * rt;
    const double C3 = l3 * (dp + rt * ct * dvn) / (2.0 * ct2);
    double d[ENF];
    d[0] = 0.5 * (C1 + C2a + C3);
    for (int k = 0; k < 3; ++k)
        d[1 + k] = 0.5 * (C1 * (ut[k] - ct * nv[k]) + C2a * ut[k] + C2b * (du[k] - dvn * nv[k]) +
                          C3 * (ut[k] + ct * nv[k]));
    d[4] = 0.5 * (C1 * (Ht - ct * vnt) + C2a * (0.5 * qt2) +
                  C2b * (ut[0] * du[0] + ut[1] * du[1] + ut[2] * du[2] - vnt * dvn) + C3 * (Ht + ct * vnt));
    for (int f = 0; f < ENF; ++f) Fh[f] = 0.5 * (nd * Fm[f] + nd * Fp[f]) + d[f];
}

__device__ __forceinline__ int eneighbour(const EulerBasis &B, int e, int ax, int side) {
    int idx[3] = {e % B.nx, (e / B.nx) % B.ny, e / (B.nx * B.ny)};
    const int dims[3] = {B.nx, B.ny, B.nz};
    idx[ax] = (idx[ax] + side + dims[ax]) % dims[ax];
    return idx[0] + B.nx * (idx[1] + B.ny * idx[2]);
}

// value of mode (i, j, k) at face point (b, c) of face (ax, side s): axis ax
// at xi = -1 / +1, the other two axes at nodes b, c in increasing axis order
__device__ __forceinline__ double eface_mode(const EulerTables &B, int ax, int s, int b, int c, const int (&id)[3]) {
    double v = 1.0;
    int t = 0;
    for (int d = 0; d < 3; ++d) {
        if (d == ax) v *= B.Lf[s][id[d]];
        else v *= B.L[id[d]][t++ == 0 ? b : c];
    }
    return v;
}

__global__ void __launch_bounds__(kEulerThreads) k_euler_residual(const EulerBasis B, const double *__restrict__ q,
                                                                  double *__restrict__ R) {
    extern __shared__ double esm[];
    __shared__ EulerTables tb;   // the basis tables, indexed at run time (LDS, not per-lane copies)
    const int e = blockIdx.x, tid = threadIdx.x, T = blockDim.x;
    for (int i = tid; i < (int)(sizeof(EulerTables) / sizeof(double)); i += T)
        reinterpret_cast<double *>(&tb)[i] = reinterpret_cast<const double *>(&B.t)[i];
    const int P1 = B.P1, nq = B.nq, nt = P1 * P1 * P1, n = ENF * nt;
    const int nv = nq * nq * nq, nf = nq * nq;
    double *cs = esm;                              // 7 * n: own, then the neighbour of face 0..5
    double *wF = cs + 7 * n;                       // [3][ENF][nv]
    double *wFh = wF + 3 * ENF * nv;               // [6][ENF][nf]
    for (int idx = tid; idx < 7 * n; idx += T) {
        const int s = idx / n, r = idx - s * n;
        const int src = s == 0 ? e : eneighbour(B, e, (s - 1) / 2, ((s - 1) & 1) ? 1 : -1);
        cs[idx] = q[(size_t)src * n + r];
    }
    __syncthreads();
    // volume points: weighted fluxes w F_d(q)
    for (int pnt = tid; pnt < nv; pnt += T) {
        const int a = pnt / (nq * nq), b = (pnt / nq) % nq, c = pnt % nq;
        double qv[ENF] = {0, 0, 0, 0, 0};
        for (int m = 0; m < nt; ++m) {
            const double ph = tb.L[m % P1][a] * tb.L[(m / P1) % P1][b] * tb.L[m / (P1 * P1)][c];
            for (int f = 0; f < ENF; ++f) qv[f] += cs[f * nt + m] * ph;
        }
        const double w = tb.w[a] * tb.w[b] * tb.w[c] * B.detJ;
        for (int d = 0; d < 3; ++d) {
            double F[ENF];
            eflux(qv, d, F);
            for (int f = 0; f < ENF; ++f) wF[(d * ENF + f) * nv + pnt] = w * F[f];
        }
    }
    // face points: weighted numerical fluxes
    for (int fp = tid; fp < 6 * nf; fp += T) {
        const int fi = fp / nf, bc = fp - fi * nf, b = bc / nq, c = bc - b * nq;
        const int ax = fi / 2, s = fi & 1;
        double qm[ENF] = {0, 0, 0, 0, 0}, qp[ENF] = {0, 0, 0, 0, 0};
        const double *cn = cs + (1 + fi) * n;
        for (int m = 0; m < nt; ++m) {
            const int id[3] = {m % P1, (m / P1) % P1, m / (P1 * P1)};
            const double phi = eface_mode(tb, ax, s, b, c, id), psi = eface_mode(tb, ax, 1 - s, b, c, id);
            for (int f = 0; f < ENF; ++f) {
                qm[f] += cs[f * nt + m] * phi;
                qp[f] += cn[f * nt + m] * psi;
            }
        }
        double Fh[ENF];
        eroe(qm, qp, ax, s ? 1.0 : -1.0, Fh);
        const double w = tb.w[b] * tb.w[c] * B.dS;
        for (int f = 0; f < ENF; ++f) wFh[(fi * ENF + f) * nf + bc] = w * Fh[f];
    }
    __syncthreads();
    // outputs (field f, mode m): - sum_volume grad(phi_m) . wF + sum_faces phi_m wFh
    for (int o = tid; o < n; o += T) {
        const int f = o / nt, m = o - f * nt;
        const int id[3] = {m % P1, (m / P1) % P1, m / (P1 * P1)};
        double vol = 0.0;
        for (int pnt = 0; pnt < nv; ++pnt) {
            const int a = pnt / (nq * nq), b = (pnt / nq) % nq, c = pnt % nq;
            const double li = tb.L[id[0]][a], lj = tb.L[id[1]][b], lk = tb.L[id[2]][c];
            const double g0 = tb.dL[id[0]][a] * lj * lk, g1 = li * tb.dL[id[1]][b] * lk, g2 = li * lj * tb.dL[id[2]][c];
            vol += (g0 * wF[(0 * ENF + f) * nv + pnt] + g1 * wF[(1 * ENF + f) * nv + pnt] +
                    g2 * wF[(2 * ENF + f) * nv + pnt]);
        }
        double fac = 0.0;
        for (int fi = 0; fi < 6; ++fi)
            for (int bc = 0; bc < nf; ++bc)
                fac += eface_mode(tb, fi / 2, fi & 1, bc / nq, bc % nq, id) * wFh[(fi * ENF + f) * nf + bc];
        R[(size_t)e * n + o] = fac - B.dxi * vol;
    }
}

// Legendre P_k and P_k' at x (k = 0..P), three-term recurrence
void elegendre(int P, double x, double *L, double *dL) {
    L[0] = 1.0;
    dL[0] = 0.0;
    if (P >= 1) { L[1] = x; dL[1] = 1.0; }
    for (int k = 2; k <= P; ++k) {
        L[k] = ((2 * k - 1) * x * L[k - 1] - (k - 1) * L[k - 2]) / k;
        dL[k] = dL[k - 2] + (2 * k - 1) * L[k - 1];
    }
}

// Gauss-Legendre nodes (ascending) and weights on [-1, 1], Newton on P_n
void egauss(int n, double *x, double *w) {
    double L[8], dL[8];
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            elegendre(n, z, L, dL);
            const double dz = L[n] / dL[n];
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        elegendre(n, z, L, dL);
        x[n - 1 - i] = z;
        w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dL[n] * dL[n]);
    }
}

}  // namespace

// R(q) on the device for the nx^3 periodic box at degree order (the layout
// and discretization of cdg_euler_residual); q, R device pointers of nrows*n.
int launch_euler_residual(cdg_ctx *ctx, int nx, int ny, int nz, int order, const double *q, double *R) {
    EulerBasis B{};
    B.P1 = order + 1;
    B.nq = order + 2;
    B.nx = nx;
    B.ny = ny;
    B.nz = nz;
    const double h = 1.0 / nx;
    B.detJ = (h / 2) * (h / 2) * (h / 2);
    B.dxi = 2.0 / h;
    B.dS = (h / 2) * (h / 2);
    double xq[5], l[4], dl[4];
    egauss(B.nq, xq, B.t.w);
    for (int a = 0; a < B.nq; ++a) {
        elegendre(order, xq[a], l, dl);
        for (int k = 0; k < B.P1; ++k) { B.t.L[k][a] = l[k]; B.t.dL[k][a] = dl[k]; }
    }
    for (int s = 0; s < 2; ++s) {
        elegendre(order, s ? 1.0 : -1.0, l, dl);
        for (int k = 0; k < B.P1; ++k) B.t.Lf[s][k] = l[k];
    }
    const int nt = B.P1 * B.P1 * B.P1, n = ENF * nt, nv = B.nq * B.nq * B.nq, nf = B.nq * B.nq;
    const size_t lds = sizeof(double) * (7 * (size_t)n + 3 * ENF * (size_t)nv + 6 * ENF * (size_t)nf);
    hipLaunchKernelGGL(k_euler_residual, dim3(nx * ny * nz), dim3(kEulerThreads), lds, ctx->stream, B, q, R);
    CDG_HIP_TRY(hipGetLastError());
    return CDG_OK;
}
