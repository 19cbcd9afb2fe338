// SU(3) heatbath and overrelaxation for the Wilson gauge action, and the
// average plaquette (physics/algorithms/su3heatbath.cpp:33-73,
// hardware/code/heatbath.cpp, gaugeobservables_plaquette.cl).
//
// Cabibbo-Marinari: each link is updated by left multiplication with SU(2)
// elements in the three subgroups (0,1), (0,2), (1,2); the SU(2) element is
// drawn with the Kennedy-Pendleton sampler (heatbath) or set to the
// microcanonical reflection (overrelaxation).  All links of one direction and
// one site parity are independent given the others, so one launch updates them
// together.  Random numbers come from a stateless hash of (seed, sweep, link,
// draw), so a configuration depends only on its seed.  Gauge fields are F64_18.
#include "cl2_device.hpp"
#include "cl2_internal.hpp"

namespace cl2 {

namespace {

__device__ __forceinline__ uint64_t hmix(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
struct Rng {
    uint64_t key, ctr;
    __device__ double uniform()  // (0, 1]
    {
        return ((hmix(key ^ hmix(ctr++)) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
    }
};

struct Coord {
    uint32_t c[4];  // t, x, y, z (direction order)
};

__device__ __forceinline__ Coord shift(const LatticeGeom& g, Coord a, int dir, int step)
{
    const uint32_t ext = dir == 0 ? g.nt : g.ns;
    a.c[dir]           = (a.c[dir] + ext + step) % ext;
    return a;
}

__device__ __forceinline__ const double2* link_ptr(const double2* U, const LatticeGeom& g, const Coord& a, int dir)
{
    const uint32_t site = a.c[1] + g.ns * (a.c[2] + g.ns * (a.c[3] + g.ns * a.c[0]));
    const uint32_t par  = (a.c[0] + a.c[1] + a.c[2] + a.c[3]) & 1u;
    return U + link_elem(dir * 2 + par, 9, 0, site >> 1, g.veo);
}

/* the 9 lanes of the link at p (tile stride LINK_TILE) */
__device__ __forceinline__ void load(const double2* p, uint32_t, Su3<double>& u)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double2 v         = p[k * LINK_TILE];
        u.e[k / 3][k % 3] = Cx<double>{v.x, v.y};
    }
}

__device__ __forceinline__ void mul(const Su3<double>& a, const Su3<double>& b, Su3<double>& r)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Cx<double> s = cmul(a.e[i][0], b.e[0][j]);
            s            = s + cmul(a.e[i][1], b.e[1][j]);
            r.e[i][j]    = s + cmul(a.e[i][2], b.e[2][j]);
        }
}
// a * b^dagger
__device__ __forceinline__ void mul_nd(const Su3<double>& a, const Su3<double>& b, Su3<double>& r)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Cx<double> s = cmulconj(a.e[i][0], b.e[j][0]);
            s            = s + cmulconj(a.e[i][1], b.e[j][1]);
            r.e[i][j]    = s + cmulconj(a.e[i][2], b.e[j][2]);
        }
}
// a^dagger * b
__device__ __forceinline__ void mul_dn(const Su3<double>& a, const Su3<double>& b, Su3<double>& r)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Cx<double> s = cmul(conj(a.e[0][i]), b.e[0][j]);
            s            = s + cmul(conj(a.e[1][i]), b.e[1][j]);
            r.e[i][j]    = s + cmul(conj(a.e[2][i]), b.e[2][j]);
        }
}

__device__ __forceinline__ void reunitarise(Su3<double>& u)
{
    double n0 = 0.;
    for (int c = 0; c < 3; ++c) n0 += u.e[0][c].re * u.e[0][c].re + u.e[0][c].im * u.e[0][c].im;
    n0 = 1.0 / sqrt(n0);
    for (int c = 0; c < 3; ++c) u.e[0][c] = scale(u.e[0][c], n0);
    Cx<double> p{0., 0.};
    for (int c = 0; c < 3; ++c) p = p + cmulconj(u.e[1][c], u.e[0][c]);
    for (int c = 0; c < 3; ++c) u.e[1][c] = u.e[1][c] - cmul(p, u.e[0][c]);
    double n1 = 0.;
    for (int c = 0; c < 3; ++c) n1 += u.e[1][c].re * u.e[1][c].re + u.e[1][c].im * u.e[1][c].im;
    n1 = 1.0 / sqrt(n1);
    for (int c = 0; c < 3; ++c) u.e[1][c] = scale(u.e[1][c], n1);
    su3_reconstruct_row2(u);
}

/* staple sum S with Re tr(U_mu(x) S) = sum of the six plaquettes containing U_mu(x) */
__device__ void staple(const double2* U, const LatticeGeom& g, const Coord& x, int mu, Su3<double>& S)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S.e[i][j] = Cx<double>{0., 0.};
    const Coord xp = shift(g, x, mu, 1);
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        Su3<double> a, b, c, t1, t2;
        // U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
        load(link_ptr(U, g, xp, nu), g.veo, a);
        load(link_ptr(U, g, shift(g, x, nu, 1), mu), g.veo, b);
        load(link_ptr(U, g, x, nu), g.veo, c);
        mul_nd(a, b, t1);
        mul_nd(t1, c, t2);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) S.e[i][j] = S.e[i][j] + t2.e[i][j];
        // U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu)
        const Coord xm = shift(g, x, nu, -1);
        load(link_ptr(U, g, shift(g, xp, nu, -1), nu), g.veo, a);
        load(link_ptr(U, g, xm, mu), g.veo, b);
        load(link_ptr(U, g, xm, nu), g.veo, c);
        mul(b, a, t1);       // U_mu(x-nu) U_nu(x+mu-nu)
        mul_dn(t1, c, t2);   // (U_mu U_nu')^dag U_nu(x-nu)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) S.e[i][j] = S.e[i][j] + t2.e[i][j];
    }
}

/* Kennedy-Pendleton: y0 with density sqrt(1 - y0^2) exp(alpha y0) on [-1, 1]; a bounded number of
 * trials, after which the caller keeps the link unchanged (probability negligible for alpha > 0.1). */
__device__ bool kp_sample(double alpha, Rng& rng, double& y0)
{
    for (int trial = 0; trial < 200; ++trial) {
        const double r1 = rng.uniform(), r2 = rng.uniform(), r3 = rng.uniform(), r4 = rng.uniform();
        const double cs = cos(6.283185307179586 * r2);
        const double d  = -(log(r1) + cs * cs * log(r3)) / alpha;
        if (r4 * r4 <= 1.0 - 0.5 * d) {
            y0 = 1.0 - d;
            return true;
        }
    }
    return false;
}

}  // namespace

/* mode 0: heatbath, 1: overrelaxation */
__global__ __launch_bounds__(128) void k_gauge_update(double2* U, LatticeGeom g, int mu, int parity, double beta,
                                                      uint64_t key, int mode)
{
    const uint32_t i = blockIdx.x * 128 + threadIdx.x;
    if (i >= g.veo) return;
    uint32_t x, y, z, t;
    site_coords(g, i, parity, x, y, z, t);
    Coord cx;
    cx.c[0] = t;
    cx.c[1] = x;
    cx.c[2] = y;
    cx.c[3] = z;
    Su3<double> S, u, R;
    staple(U, g, cx, mu, S);
    double2* up = const_cast<double2*>(link_ptr(U, g, cx, mu));
    load(up, g.veo, u);
    Rng rng{hmix(key ^ ((uint64_t)mu << 56) ^ ((uint64_t)i << 4)), 0};
    const int sub[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int s = 0; s < 3; ++s) {
        const int a = sub[s][0], b = sub[s][1];
        mul(u, S, R);
        // quaternion components of the 2x2 block: R_sub = k V, V = v0 + i v.sigma
        double v0 = 0.5 * (R.e[a][a].re + R.e[b][b].re);
        double v3 = 0.5 * (R.e[a][a].im - R.e[b][b].im);
        double v1 = 0.5 * (R.e[a][b].im + R.e[b][a].im);
        double v2 = 0.5 * (R.e[a][b].re - R.e[b][a].re);
        const double k = sqrt(v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3);
        if (k < 1e-300) continue;
        v0 /= k, v1 /= k, v2 /= k, v3 /= k;
        // X = Y V^dagger (heatbath) or X = V^dagger V^dagger (overrelaxation), as quaternions q = (q0, q1, q2, q3)
        double y0, y1, y2, y3;
        if (mode == 0) {
            if (!kp_sample(2.0 * beta * k / 3.0, rng, y0)) continue;
            const double rad = sqrt(fmax(0.0, 1.0 - y0 * y0));
            const double ct  = 2.0 * rng.uniform() - 1.0;
            const double st  = sqrt(fmax(0.0, 1.0 - ct * ct));
            const double ph  = 6.283185307179586 * rng.uniform();
            y1 = rad * st * cos(ph);
            y2 = rad * st * sin(ph);
            y3 = rad * ct;
        } else {
            y0 = v0, y1 = -v1, y2 = -v2, y3 = -v3;  // Y = V^dagger, so X = V^dagger V^dagger
        }
        // quaternion product X = Y * conj(V), conj(V) = (v0, -v1, -v2, -v3)
        const double w0 = v0, w1 = -v1, w2 = -v2, w3 = -v3;
        const double x0 = y0 * w0 - y1 * w1 - y2 * w2 - y3 * w3;
        const double x1 = y0 * w1 + y1 * w0 - y2 * w3 + y3 * w2;
        const double x2 = y0 * w2 + y2 * w0 - y3 * w1 + y1 * w3;
        const double x3 = y0 * w3 + y3 * w0 - y1 * w2 + y2 * w1;
        // X as 2x2: [[x0 + i x3, x2 + i x1], [-x2 + i x1, x0 - i x3]], acting on rows a, b
        const Cx<double> X00{x0, x3}, X01{x2, x1}, X10{-x2, x1}, X11{x0, -x3};
        for (int c = 0; c < 3; ++c) {
            const Cx<double> ua = u.e[a][c], ub = u.e[b][c];
            u.e[a][c]           = cmul(X00, ua) + cmul(X01, ub);
            u.e[b][c]           = cmul(X10, ua) + cmul(X11, ub);
        }
    }
    reunitarise(u);
#pragma unroll
    for (int q = 0; q < 9; ++q) up[q * LINK_TILE] = make_double2(u.e[q / 3][q % 3].re, u.e[q / 3][q % 3].im);
}

/* per-block sum over sites of sum_{mu<nu} Re tr P_mu,nu(x); grid-stride over both parities */
__global__ __launch_bounds__(256) void k_plaquette(const double2* U, LatticeGeom g, double* partials)
{
    __shared__ double lds[8];
    double acc = 0.;
    for (uint32_t l = blockIdx.x * 256 + threadIdx.x; l < 2 * g.veo; l += gridDim.x * 256) {
        const int parity = l >= g.veo ? 1 : 0;
        const uint32_t i = l - (parity ? g.veo : 0);
        uint32_t x, y, z, t;
        site_coords(g, i, parity, x, y, z, t);
        Coord cx;
        cx.c[0] = t;
        cx.c[1] = x;
        cx.c[2] = y;
        cx.c[3] = z;
        for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu) {
                Su3<double> a, b, c, d, t1, t2;
                load(link_ptr(U, g, cx, mu), g.veo, a);
                load(link_ptr(U, g, shift(g, cx, mu, 1), nu), g.veo, b);
                load(link_ptr(U, g, shift(g, cx, nu, 1), mu), g.veo, c);
                load(link_ptr(U, g, cx, nu), g.veo, d);
                mul(a, b, t1);
                mul_nd(t1, c, t2);
                // Re tr(t2 d^dagger)
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) acc += t2.e[p][q].re * d.e[p][q].re + t2.e[p][q].im * d.e[p][q].im;
            }
    }
    const double s = block_sum<256>(acc, lds);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = s, partials[2 * blockIdx.x + 1] = 0.;
}

hipError_t launch_gauge_update(cl2_ctx* c, void* gauge, double beta, int mode, uint64_t key)
{
    const unsigned nb = blocks_for(c->g.veo, 128);
    for (int mu = 0; mu < 4; ++mu)
        for (int par = 0; par < 2; ++par) {
            hipLaunchKernelGGL(k_gauge_update, dim3(nb), dim3(128), 0, c->stream, reinterpret_cast<double2*>(gauge),
                               c->g, mu, par, beta, key * 8 + mu * 2 + par, mode);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

hipError_t launch_plaquette(cl2_ctx* c, const void* gauge, double* res)
{
    const unsigned nb = (unsigned)std::min<size_t>(blocks_for(2ull * c->g.veo, 256), REDUCE_MAX_BLOCKS);
    hipLaunchKernelGGL(k_plaquette, dim3(nb), dim3(256), 0, c->stream, reinterpret_cast<const double2*>(gauge), c->g,
                       c->partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_stage2(c, nb, res, 0, nullptr, 0);
}

}  // namespace cl2
