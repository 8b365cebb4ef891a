// Monte-Carlo cycles of the two-species RuNNer potential in the NNVT
// ensemble, compiled once per element type N2R (double: FP64 throughout;
// float: binary32 positions, distances, features and networks, FP64 energy
// sums and Metropolis test) by ptmc_nnp2_mc.hip and ptmc_nnp2_mc_f32.hip.
// Reference: RuNNerPotential2Atom (src/EnergyEvaluation.jl:1287-1448,
// :1622-1643, :1748-1768, :1824-1856), NNVT (src/Ensembles.jl:138-150,
// :270-282), moves src/MCMoves.jl:115-145, :363-385, commit
// src/swap_config.jl:223-277.
//
// One workgroup owns one configuration for a cycle of N + n_swaps moves, as
// the single-species kernels do.  Its positions, squared distances and
// cutoffs (R2, FC) and the trial feature matrix D live in LDS; the current
// feature matrix G and per-atom energies stay in the slot's HBM buffers.
// A move is an atom displacement or, with probability n_swaps / (N +
// n_swaps), an exchange of one atom of each species (slot semantics: the
// atoms keep their species and trade positions).  For the moved set M (one
// atom, or the two swapped ones) the trial features are
//   * columns of M: rebuilt in full from the trial positions, the partner
//     pairs of an angular sum split into NNP2_NCH chunks whose partial sums
//     are reduced in a fixed order;
//   * columns within the cutoff of an old or new position of M: G plus the
//     change of every radial pair and angular triangle holding an atom of M
//     (new minus old, per-species scale);
//   * other columns: unchanged, not built.
// Only changed columns are built; the networks run on them alone (an
// unchanged column gives the same atomic energy bit for bit), the
// Metropolis test uses the FP64 energy change, and an accepted move writes
// the changed columns back.  The reference builds the same trial matrix as
// G plus deltas accumulated per pair / triangle (DeltaMatrix.jl:23-624);
// every sum here is gathered by one thread in a fixed order, so a run does
// not depend on the schedule and agrees with the reference's arithmetic to
// rounding.  Angular functions of equal eta form groups that share each
// triangle's cosine and exponential; (column, group) and (moved column,
// group, chunk) tasks come from an LDS queue, heaviest first.
// (no include guard: included once per element type)
#include "ptmc_nnp_layout.hpp"
#include "ptmc_sweep.hpp"

namespace ptmc {

namespace {

struct N2Lds {
    N2R *X, *XT, *R2, *FC, *RN, *FN, *D, *eat, *eatn, *W, *S;
    int *mi, *chg;
};

__device__ inline N2Lds n2_lds(unsigned char *raw, int N, int ng, int nfn, int h1, int h2)
{
    N2Lds l;
    N2R *p = reinterpret_cast<N2R *>(raw);
    l.X = p;
    l.XT = l.X + 3 * N;
    l.R2 = l.XT + 3 * N;
    l.FC = l.R2 + N * N;
    l.RN = l.FC + N * N;
    l.FN = l.RN + 2 * N;
    l.D = l.FN + 2 * N;
    l.eat = l.D + (size_t)N * ng;
    l.eatn = l.eat + N;
    l.W = l.eatn + N;
    l.S = l.W + 2 * nnp2_params(ng, h1, h2);
    l.mi = reinterpret_cast<int *>(l.S + nnp2_scratch_doubles(N, nfn, h1, h2));
    l.chg = l.mi + N;
    return l;
}

// type-2 cutoff (MachineLearningPotential/Cutoff.jl:79-90)
__device__ __forceinline__ N2R n2_cut(N2R r, N2R rc)
{
    return r < rc ? N2R(0.5) * (cos(N2R(3.14159265358979323846) * (r / rc)) + N2R(1.0)) : N2R(0.0);
}

__device__ __forceinline__ N2R n2_d2(const N2R *a, const N2R *b)
{
    const N2R dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return dx * dx + dy * dy + dz * dz;
}

// radial pair value f exp(-eta r^2), zero at f = 0 or 1 (SymmFunc.jl:229-231)
__device__ __forceinline__ N2R n2_rad(N2R r2, N2R f, N2R eta)
{
    return (f != N2R(0.0) && f != N2R(1.0)) ? f * exp(-eta * r2) : N2R(0.0);
}

// (1 + lambda cos)^zeta: integer zeta by squaring, else pow
__device__ __forceinline__ N2R n2_pow(N2R b, N2R zeta)
{
    if (zeta >= N2R(0.0) && zeta <= N2R(64.0) && zeta == rint(zeta)) {
        int e = (int)zeta;
        N2R r = N2R(1.0), q = b;
        while (e) {
            if (e & 1) r *= q;
            q *= q;
            e >>= 1;
        }
        return r;
    }
    return pow(b, zeta);
}

// squared distance / cutoff between atoms a and b in the trial configuration
__device__ __forceinline__ N2R n2_rt(const N2Lds &L, int N, int a, int b)
{
    const int ma = L.mi[a], mb = L.mi[b];
    return ma >= 0 ? L.RN[ma * N + b] : (mb >= 0 ? L.RN[mb * N + a] : L.R2[a * N + b]);
}
__device__ __forceinline__ N2R n2_ft(const N2Lds &L, int N, int a, int b)
{
    const int ma = L.mi[a], mb = L.mi[b];
    return ma >= 0 ? L.FN[ma * N + b] : (mb >= 0 ? L.FN[mb * N + a] : L.FC[a * N + b]);
}

// R2 / FC of the configuration X
__device__ void n2_tables(const N2Lds &L, int N, N2R rc)
{
    for (int e = threadIdx.x; e < N * N; e += NNP2_NT) {
        const int p = e / N, q = e % N;
        const N2R r2 = p == q ? N2R(0.0) : n2_d2(&L.X[3 * p], &L.X[3 * q]);
        L.R2[e] = r2;
        L.FC[e] = n2_cut(sqrt(r2), rc);
    }
}

// An angular group: the functions of one eta (at most NNP2_GMAX), which
// share every triangle's geometry and exponential.
struct N2Grp {
    int cnt;
    int q[NNP2_GMAX];
    N2R lam[NNP2_GMAX], zeta[NNP2_GMAX];
    N2R eta;
};

__device__ __forceinline__ void n2_load_group(const Nnp2Tables &T, int g, N2Grp &G)
{
    const int first = T.grp[2 * g], cnt = T.grp[2 * g + 1];
    G.cnt = cnt;
#pragma unroll
    for (int l = 0; l < NNP2_GMAX; ++l) {
        const int q = T.gfn[first + (l < cnt ? l : 0)];
        G.q[l] = q;
        G.lam[l] = T.ang[(size_t)q * 15 + 1];
        G.zeta[l] = T.ang[(size_t)q * 15 + 2];
    }
    G.eta = T.ang[(size_t)T.gfn[first] * 15];
}

// cosine at c and exp(-eta (r_cj^2 + r_ck^2 + r_jk^2)) f_cj f_ck f_jk of triangle (c; j, k)
// (SymmFunc.jl:184-202, Cutoff.jl:48-68)
__device__ __forceinline__ void n2_tri_geom(const N2R *pc, const N2R *pj, const N2R *pk, N2R rcj,
                                            N2R rck, N2R rjk, N2R fcj, N2R fck, N2R fjk, N2R eta,
                                            N2R &cth, N2R &e)
{
    const N2R ux = pj[0] - pc[0], uy = pj[1] - pc[1], uz = pj[2] - pc[2];
    const N2R wx = pk[0] - pc[0], wy = pk[1] - pc[1], wz = pk[2] - pc[2];
    cth = (ux * wx + uy * wy + uz * wz) / (sqrt(rcj) * sqrt(rck));
    e = exp(-eta * (rcj + rck + rjk)) * fcj * fck * fjk;
}

// acc[l] += e (1 + lambda_l cos)^zeta_l for the functions of the group
__device__ __forceinline__ void n2_group_add(const N2Grp &G, N2R cth, N2R e, N2R (&acc)[NNP2_GMAX])
{
#pragma unroll
    for (int l = 0; l < NNP2_GMAX; ++l)
        if (l < G.cnt) acc[l] += e * n2_pow(N2R(1.0) + G.lam[l] * cth, G.zeta[l]);
}

// Radial sums of function q at centre c over the trial configuration:
// s[0] partners of species 1, s[1] of species 2.
__device__ void n2_radial_sums(const Nnp2Tables &T, const N2Lds &L, int N, int c, int q, N2R s[2])
{
    const int n1 = T.n1;
    const N2R eta = (N2R)T.rad[(size_t)q * 9];
    s[0] = s[1] = N2R(0.0);
    for (int j = 0; j < N; ++j) {
        if (j == c) continue;
        s[j < n1 ? 0 : 1] += n2_rad(n2_rt(L, N, c, j), n2_ft(L, N, c, j), eta);
    }
}

// Angular sums of group g at centre c over the trial configuration XT, over
// the partner pairs (j, k) with j = ch, ch + nch, ... (strided, so the
// chunks mix near and far partners): out(l, row, v) adds v to row `row`
// (partner pair 11 / 12 / 22) of the group's function l.  b is the other
// moved atom of a swap (mb its row of RN / FN) or -1; the centre's own trial
// row is read directly.
template <typename Out>
__device__ __forceinline__ void n2_group_sums(const Nnp2Tables &T, const N2Lds &L, const N2Grp &G, int N, int c, int b,
                                              int mb, int ch, int nch, Out out)
{
    const int n1 = T.n1;
    const int mc = L.mi[c];
    const N2R *rc_row = mc >= 0 ? &L.RN[mc * N] : &L.R2[c * N];
    const N2R *fc_row = mc >= 0 ? &L.FN[mc * N] : &L.FC[c * N];
    const N2R *pc = &L.XT[3 * c];
    for (int j = ch; j < N; j += nch) {
        if (j == c) continue;
        const N2R fcj = fc_row[j];
        if (fcj == N2R(0.0)) continue;
        const N2R rcj = rc_row[j];
        const int sj = j < n1 ? 0 : 1;
        for (int blk = 0; blk < 2; ++blk) {
            const int klo = blk ? (j + 1 > n1 ? j + 1 : n1) : j + 1, khi = blk ? N : n1;
            if (klo >= khi) continue;
            N2R acc[NNP2_GMAX];
#pragma unroll
            for (int l = 0; l < NNP2_GMAX; ++l) acc[l] = N2R(0.0);
            for (int k = klo; k < khi; ++k) {
                if (k == c) continue;
                const N2R fck = fc_row[k];
                if (fck == N2R(0.0)) continue;
                const N2R fjk = j == b ? L.FN[mb * N + k] : (k == b ? L.FN[mb * N + j] : L.FC[j * N + k]);
                if (fjk == N2R(0.0)) continue;
                const N2R rjk = j == b ? L.RN[mb * N + k] : (k == b ? L.RN[mb * N + j] : L.R2[j * N + k]);
                N2R cth, e;
                n2_tri_geom(pc, &L.XT[3 * j], &L.XT[3 * k], rcj, rc_row[k], rjk, fcj, fck, fjk, G.eta, cth, e);
                n2_group_add(G, cth, e, acc);
            }
#pragma unroll
            for (int l = 0; l < NNP2_GMAX; ++l)
                if (l < G.cnt) out(l, sj + blk, acc[l]);
        }
    }
}

// scaled rows of radial function q / of the group at a centre of species sc
__device__ __forceinline__ void n2_store_radial(const Nnp2Tables &T, N2R *Dc, int q, int sc, const N2R s[2])
{
    const double *R = T.rad + (size_t)q * 9;
    for (int r = 0; r < 2; ++r) Dc[2 * q + r] = s[r] * R[1 + 2 * sc + r] + R[5 + 2 * sc + r];
}
// radial rows of function q at an unmoved centre c: G plus the change of
// every pair with a moved atom (ma[0..nm-1]), per-species scale
__device__ void n2_delta_radial(const Nnp2Tables &T, const N2Lds &L, const N2R *Gc, int N, int c, int q, int nm,
                                const int *ma)
{
    const int n1 = T.n1, sc = c < n1 ? 0 : 1;
    const double *R = T.rad + (size_t)q * 9;
    const N2R eta = R[0];
    N2R acc[2] = {N2R(0.0), N2R(0.0)};
    for (int m = 0; m < nm; ++m) {
        const int a = ma[m];
        acc[a < n1 ? 0 : 1] +=
            n2_rad(L.RN[m * N + c], L.FN[m * N + c], eta) - n2_rad(L.R2[c * N + a], L.FC[c * N + a], eta);
    }
    N2R *Dc = &L.D[(size_t)c * T.ng];
    for (int r = 0; r < 2; ++r) Dc[2 * q + r] = Gc[2 * q + r] + acc[r] * R[1 + 2 * sc + r];
}

// angular rows of group g at an unmoved centre c: G plus the change of every
// triangle with a moved atom (new minus old), per-species scale, folded into
// D block by block
__device__ __forceinline__ void n2_delta_group(const Nnp2Tables &T, const N2Lds &L, const N2Grp &G, const N2R *Gc,
                                               int N, int c, int nm, const int *ma)
{
    const int n1 = T.n1, sc = c < n1 ? 0 : 1;
    const N2R *pc = &L.X[3 * c];
    N2R *Dc = &L.D[(size_t)c * T.ng];
#pragma unroll
    for (int l = 0; l < NNP2_GMAX; ++l) {
        if (l >= G.cnt) continue;
        const int r = 2 * T.n_rad + 3 * G.q[l];
        Dc[r] = Gc[r];
        Dc[r + 1] = Gc[r + 1];
        Dc[r + 2] = Gc[r + 2];
    }
    for (int m = 0; m < nm; ++m) {
        const int a = ma[m], sa = a < n1 ? 0 : 1;
        const N2R foca = L.FC[c * N + a], fnca = L.FN[m * N + c];
        if (foca == N2R(0.0) && fnca == N2R(0.0)) continue;
        const N2R roca = L.R2[c * N + a], rnca = L.RN[m * N + c];
        for (int blk = 0; blk < 2; ++blk) {
            N2R acc[NNP2_GMAX];
#pragma unroll
            for (int l = 0; l < NNP2_GMAX; ++l) acc[l] = N2R(0.0);
            const int klo = blk ? n1 : 0, khi = blk ? N : n1;
            for (int k = klo; k < khi; ++k) {
                if (k == c || k == a) continue;
                const int mk = L.mi[k];
                if (mk >= 0 && mk < m) continue;  // the pair of two moved atoms counts once
                N2R cth, e;
                const N2R fck = L.FC[c * N + k];
                if (foca != N2R(0.0) && fck != N2R(0.0)) {
                    const N2R fak = L.FC[a * N + k];
                    if (fak != N2R(0.0)) {
                        n2_tri_geom(pc, &L.X[3 * a], &L.X[3 * k], roca, L.R2[c * N + k], L.R2[a * N + k], foca, fck,
                                    fak, G.eta, cth, e);
                        n2_group_add(G, cth, -e, acc);
                    }
                }
                const N2R fnck = mk >= 0 ? L.FN[mk * N + c] : fck;
                if (fnca != N2R(0.0) && fnck != N2R(0.0)) {
                    const N2R fnak = L.FN[m * N + k];
                    if (fnak != N2R(0.0)) {
                        const N2R rnck = mk >= 0 ? L.RN[mk * N + c] : L.R2[c * N + k];
                        n2_tri_geom(pc, &L.XT[3 * a], &L.XT[3 * k], rnca, rnck, L.RN[m * N + k], fnca, fnck, fnak,
                                    G.eta, cth, e);
                        n2_group_add(G, cth, e, acc);
                    }
                }
            }
            const int row = sa + blk;
#pragma unroll
            for (int l = 0; l < NNP2_GMAX; ++l) {
                if (l >= G.cnt) continue;
                const double *A = T.ang + (size_t)G.q[l] * 15;
                Dc[2 * T.n_rad + 3 * G.q[l] + row] += acc[l] * A[3 + 3 * sc + row];
            }
        }
    }
}

__device__ __forceinline__ N2R n2_act(int code, N2R s) { return code == 2 ? tanh(s) : s; }

// per-atom energies of the listed columns of D (RuNNer layout: per layer the
// input-major weights, then the biases; one network per species)
__device__ void n2_forward(const Nnp2Tables &T, const N2Lds &L, int N, int nc, const int *list)
{
    const int tid = threadIdx.x, ng = T.ng, h1 = T.h1, h2 = T.h2, n1 = T.n1;
    const size_t np = nnp2_params(ng, h1, h2);
    N2R *H1 = L.S, *H2 = L.S + (size_t)N * h1;
    for (int t = tid; t < nc * h1; t += NNP2_NT) {
        const int ci = t / h1, j = t % h1, c = list[ci];
        const N2R *p = c < n1 ? L.W : L.W + np;
        const N2R *Dc = &L.D[(size_t)c * ng];
        N2R s = N2R(0.0);
        for (int k = 0; k < ng; ++k) s += Dc[k] * p[(size_t)k * h1 + j];
        s += p[(size_t)ng * h1 + j];
        H1[ci * h1 + j] = n2_act(T.act1, s);
    }
    __syncthreads();
    const size_t o2 = (size_t)ng * h1 + h1;
    for (int t = tid; t < nc * h2; t += NNP2_NT) {
        const int ci = t / h2, j = t % h2, c = list[ci];
        const N2R *p = (c < n1 ? L.W : L.W + np) + o2;
        N2R s = N2R(0.0);
        for (int k = 0; k < h1; ++k) s += H1[ci * h1 + k] * p[(size_t)k * h2 + j];
        s += p[(size_t)h1 * h2 + j];
        H2[ci * h2 + j] = n2_act(T.act2, s);
    }
    __syncthreads();
    const size_t o3 = o2 + (size_t)h1 * h2 + h2;
    for (int ci = tid; ci < nc; ci += NNP2_NT) {
        const int c = list[ci];
        const N2R *p = (c < n1 ? L.W : L.W + np) + o3;
        N2R s = N2R(0.0);
        for (int k = 0; k < h2; ++k) s += H2[ci * h2 + k] * p[k];
        s += p[h2];
        L.eatn[c] = n2_act(T.act3, s);
    }
    __syncthreads();
}

// every column rebuilt from X (XT = X, nothing moved): D, eatn and the total
__device__ double n2_rebuild(const Nnp2Tables &T, const N2Lds &L, int N, double *s_tot)
{
    for (int e = threadIdx.x; e < 3 * N; e += NNP2_NT) L.XT[e] = L.X[e];
    __syncthreads();
    const int nga = T.n_groups, nt = N * (nga + T.n_rad);
    for (int t = threadIdx.x; t < nt; t += NNP2_NT) {
        N2R *Dc;
        if (t < N * nga) {
            const int c = t / nga, g = t % nga;
            N2Grp G;
            n2_load_group(T, g, G);
            Dc = &L.D[(size_t)c * T.ng];
            const int sc = c < T.n1 ? 0 : 1;
            // rows start at the offsets; each block adds its scaled sum
#pragma unroll
            for (int l = 0; l < NNP2_GMAX; ++l) {
                if (l >= G.cnt) continue;
                const double *A = T.ang + (size_t)G.q[l] * 15;
                for (int r = 0; r < 3; ++r) Dc[2 * T.n_rad + 3 * G.q[l] + r] = A[9 + 3 * sc + r];
            }
            n2_group_sums(T, L, G, N, c, -1, 0, 0, 1, [&](int l, int row, N2R v) {
                const double *A = T.ang + (size_t)G.q[l] * 15;
                Dc[2 * T.n_rad + 3 * G.q[l] + row] += v * A[3 + 3 * sc + row];
            });
        } else {
            const int u = t - N * nga, c = u / T.n_rad, q = u % T.n_rad;
            N2R s[2];
            n2_radial_sums(T, L, N, c, q, s);
            Dc = &L.D[(size_t)c * T.ng];
            n2_store_radial(T, Dc, q, c < T.n1 ? 0 : 1, s);
        }
    }
    for (int c = threadIdx.x; c < N; c += NNP2_NT) L.chg[c] = c;
    __syncthreads();
    n2_forward(T, L, N, N, L.chg);
    if (threadIdx.x == 0) {
        double tot = 0.0;
        for (int c = 0; c < N; ++c) tot += (double)L.eatn[c];
        *s_tot = tot;
    }
    __syncthreads();
    return *s_tot;
}

__device__ void n2_load_x(const SweepArgs &a, const N2Lds &L, int slot, int N)
{
    const N2R *gX = reinterpret_cast<const N2R *>(a.coords) + (size_t)slot * 3 * NNP_NP;
    for (int e = threadIdx.x; e < 3 * N; e += NNP2_NT) {
        const int j = e / 3, c = e % 3;
        L.X[e] = gX[c * NNP_NP + j];
        L.XT[e] = L.X[e];
    }
    for (int e = threadIdx.x; e < N; e += NNP2_NT) L.mi[e] = -1;
    // both networks, staged once per launch
    const Nnp2Tables &T = a.nnp2;
    const int np = (int)nnp2_params(T.ng, T.h1, T.h2);
    for (int e = threadIdx.x; e < 2 * np; e += NNP2_NT) L.W[e] = e < np ? T.p1[e] : T.p2[e - np];
}

// set_variables + initialise_energy (src/EnergyEvaluation.jl:1824-1856, :1748-1768)
__global__ void __launch_bounds__(NNP2_NT) nnp2_init_kernel(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double s_tot;
    const int slot = blockIdx.x;
    if (a.init_run >= 0 && a.slot_run[slot] != a.init_run) return;  // whole workgroup
    const Nnp2Tables &T = a.nnp2;
    const int N = a.runs[a.slot_run[slot]].n_atoms;
    const N2Lds L = n2_lds(lds_raw, N, T.ng, T.n_rad + T.n_ang, T.h1, T.h2);
    n2_load_x(a, L, slot, N);
    __syncthreads();
    n2_tables(L, N, T.r_cut);
    __syncthreads();
    const double E = n2_rebuild(T, L, N, &s_tot);
    N2R *gG = reinterpret_cast<N2R *>(T.feat) + (size_t)slot * NNP_NP * T.ng;
    for (int e = threadIdx.x; e < N * T.ng; e += NNP2_NT) gG[e] = L.D[e];
    for (int e = threadIdx.x; e < NNP_NP; e += NNP2_NT) reinterpret_cast<N2R *>(T.eatom)[(size_t)slot * NNP_NP + e] = e < N ? L.eatn[e] : N2R(0.0);
    if (threadIdx.x == 0) a.slot_en[slot] = E;
}

// One cycle (N + n_swaps moves) of one configuration per workgroup
// (mc_step! / mc_move!, src/MCRun.jl:142-182, with get_energy!(NNVT) :87-112).
__global__ void __launch_bounds__(NNP2_NT) N2_SWEEP_ATTR nnp2_sweep_kernel(SweepArgs a)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int s_err, s_nc, s_next;
    __shared__ double s_en;
    const Nnp2Tables &T = a.nnp2;
    const int slot = blockIdx.x, tid = threadIdx.x;
    const RunInfo &ri = a.runs[a.slot_run[slot]];
    const int N = ri.n_atoms, n1 = T.n1, ng = T.ng, nfn = T.n_rad + T.n_ang;
    const uint32_t nsteps = (uint32_t)(N + T.n_swaps);
    const int t = a.slot_temp[slot];
    const int Tg = ri.temp_base + t;
    const uint32_t sweep = (uint32_t)a.sweep0;
    const double beta = a.t_beta[Tg];
    const N2R delta_r = (N2R)a.t_displ[2 * Tg];
    const N2R radius2 = (N2R)ri.radius2;
    const N2Lds L = n2_lds(lds_raw, N, ng, nfn, T.h1, T.h2);
    N2R *gG = reinterpret_cast<N2R *>(T.feat) + (size_t)slot * NNP_NP * ng;

    n2_load_x(a, L, slot, N);
    for (int e = tid; e < N; e += NNP2_NT) L.eat[e] = reinterpret_cast<const N2R *>(T.eatom)[(size_t)slot * NNP_NP + e];
    if (tid == 0) s_err = 0;
    double E = a.slot_en[slot];
    __syncthreads();
    n2_tables(L, N, T.r_cut);
    __syncthreads();
    long long nacc = 0, nsw = 0, nsw_acc = 0;

    for (uint32_t m = 0; m < nsteps; ++m) {
        // draws (every thread the same): the move index as mc_move! draws it
        // (src/MCRun.jl:145), then the trial position (src/MCMoves.jl:115-135)
        // or one atom of each species (src/MCMoves.jl:363-370)
        const MoveDraw<N2R> d0 = move_draw<N2R>(a.seed, ri.run_key, (uint32_t)t, sweep, m, 0u, nsteps);
        const bool swap = d0.index >= (uint32_t)N;
        int ma[2] = {(int)d0.index, -1};
        N2R tx = N2R(0.0), ty = N2R(0.0), tz = N2R(0.0);
        if (swap) {
            const u32x4 S = draw(a.seed, ri.run_key, RNG_MOVE, 254u, (uint32_t)t, sweep, m);
            ma[0] = (int)below(S.x, (uint32_t)n1);
            ma[1] = n1 + (int)below(S.y, (uint32_t)(N - n1));
        } else {
            const int i = ma[0];
            make_trial<N2R, BC_SPHERE, false>(L.X[3 * i], L.X[3 * i + 1], L.X[3 * i + 2], d0, delta_r, N2R(0.0), N2R(0.0),
                                                 tx, ty, tz);
            int count = 0;
            uint32_t attempt = 0;
            while (outside_sphere<N2R>(tx, ty, tz, radius2)) {
                count += 1;
                if (count == 50) {
                    // recentre! (src/Configurations.jl:79-85)
                    __syncthreads();
                    if (tid == 0) {
                        double cm[3] = {0.0, 0.0, 0.0};
                        for (int b = 0; b < N; ++b)
                            for (int k = 0; k < 3; ++k) cm[k] += L.X[3 * b + k];
                        for (int k = 0; k < 3; ++k) cm[k] /= (double)N;
                        for (int b = 0; b < N; ++b)
                            for (int k = 0; k < 3; ++k) L.X[3 * b + k] = (N2R)((double)L.X[3 * b + k] - cm[k]);
                    }
                    __syncthreads();
                } else {
                    attempt += 1;
                    const MoveDraw<N2R> dd =
                        move_draw<N2R>(a.seed, ri.run_key, (uint32_t)t, sweep, m, attempt, nsteps);
                    make_trial<N2R, BC_SPHERE, false>(L.X[3 * i], L.X[3 * i + 1], L.X[3 * i + 2], dd, delta_r,
                                                         N2R(0.0), N2R(0.0), tx, ty, tz);
                    if (count == 100) {
                        if (tid == 0) s_err = 1;
                        break;
                    }
                }
            }
        }
        const int nm = swap ? 2 : 1;
        // trial positions and the moved set
        for (int e = tid; e < 3 * N; e += NNP2_NT) {
            const int j = e / 3, k = e % 3;
            N2R v = L.X[e];
            if (!swap && j == ma[0]) v = k == 0 ? tx : (k == 1 ? ty : tz);
            if (swap && j == ma[0]) v = L.X[3 * ma[1] + k];
            if (swap && j == ma[1]) v = L.X[3 * ma[0] + k];
            L.XT[e] = v;
        }
        if (tid == 0) {
            L.mi[ma[0]] = 0;
            if (swap) L.mi[ma[1]] = 1;
            s_next = NNP2_NT;
        }
        __syncthreads();
        // distances and cutoffs of the moved atoms in the trial configuration
        for (int e = tid; e < nm * N; e += NNP2_NT) {
            const int mm = e / N, k = e % N, aa = ma[mm];
            const N2R r2 = k == aa ? N2R(0.0) : n2_d2(&L.XT[3 * aa], &L.XT[3 * k]);
            L.RN[e] = r2;
            L.FN[e] = n2_cut(sqrt(r2), T.r_cut);
        }
        __syncthreads();
        // changed columns: the moved atoms and every atom within the cutoff
        // of an old or new position of one
        auto changed = [&](int c) {
            if (L.mi[c] >= 0) return true;
            for (int mm = 0; mm < nm; ++mm)
                if (L.FC[c * N + ma[mm]] != N2R(0.0) || L.FN[mm * N + c] != N2R(0.0)) return true;
            return false;
        };
        if (tid < 64) {
            const bool f = tid < N && changed(tid);
            const uint64_t b = __ballot(f);
            if (f) L.chg[__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u))] = tid;
            if (tid == 0) s_nc = __builtin_popcountll(b);
        }
        // trial features from a queue of tasks, heaviest first: angular groups
        // of unmoved changed columns (delta) and chunks of the moved columns
        // (partial sums), then the radial functions of both
        const int nga = T.n_groups, nAg = N * nga, nBg = nm * nga * NNP2_NCH, nAr = N * T.n_rad;
        const int ntask = nAg + nBg + nAr + nm * T.n_rad;
        for (int u = tid; u < ntask; u = atomicAdd(&s_next, 1)) {
            if (u < nAg) {
                // group-major: the lanes of a wave mostly share one group
                const int g = u / N, c = u % N;
                if (L.mi[c] >= 0 || !changed(c)) continue;  // moved: rebuilt below; unchanged: not read
                N2Grp G;
                n2_load_group(T, g, G);
                n2_delta_group(T, L, G, gG + (size_t)c * ng, N, c, nm, ma);
            } else if (u < nAg + nBg) {
                const int v = u - nAg, mm = v / (nga * NNP2_NCH), g = (v / NNP2_NCH) % nga, ch = v % NNP2_NCH;
                N2Grp G;
                n2_load_group(T, g, G);
#pragma unroll
                for (int l = 0; l < NNP2_GMAX; ++l) {
                    if (l >= G.cnt) continue;
                    N2R *P = &L.S[(((size_t)mm * nfn + T.n_rad + G.q[l]) * NNP2_NCH + ch) * 3];
                    P[0] = P[1] = P[2] = N2R(0.0);
                }
                const int other = nm == 2 ? ma[1 - mm] : -1;
                n2_group_sums(T, L, G, N, ma[mm], other, 1 - mm, ch, NNP2_NCH, [&](int l, int row, N2R v) {
                    L.S[(((size_t)mm * nfn + T.n_rad + G.q[l]) * NNP2_NCH + ch) * 3 + row] += v;
                });
            } else if (u < nAg + nBg + nAr) {
                const int v = u - nAg - nBg, c = v / T.n_rad, q = v % T.n_rad;
                if (L.mi[c] >= 0 || !changed(c)) continue;
                n2_delta_radial(T, L, gG + (size_t)c * ng, N, c, q, nm, ma);
            } else {
                const int v = u - nAg - nBg - nAr, mm = v / T.n_rad, q = v % T.n_rad;
                N2R s2[2];
                n2_radial_sums(T, L, N, ma[mm], q, s2);
                N2R *P = &L.S[((size_t)mm * nfn + q) * NNP2_NCH * 3];
                P[0] = s2[0];
                P[1] = s2[1];
                P[2] = N2R(0.0);
                for (int e = 3; e < 3 * NNP2_NCH; ++e) P[e] = N2R(0.0);
            }
        }
        __syncthreads();
        for (int u = tid; u < nm * nfn; u += NNP2_NT) {
            const int mm = u / nfn, fn = u % nfn, c = ma[mm];
            N2R sum[3] = {N2R(0.0), N2R(0.0), N2R(0.0)};
            const N2R *P = &L.S[((size_t)mm * nfn + fn) * NNP2_NCH * 3];
            for (int ch = 0; ch < NNP2_NCH; ++ch)
                for (int r = 0; r < 3; ++r) sum[r] += P[3 * ch + r];
            N2R *Dc = &L.D[(size_t)c * ng];
            const int sc = c < n1 ? 0 : 1;
            if (fn < T.n_rad) {
                n2_store_radial(T, Dc, fn, sc, sum);
            } else {
                const double *A = T.ang + (size_t)(fn - T.n_rad) * 15;
                for (int r = 0; r < 3; ++r)
                    Dc[2 * T.n_rad + 3 * (fn - T.n_rad) + r] = sum[r] * A[3 + 3 * sc + r] + A[9 + 3 * sc + r];
            }
        }
        __syncthreads();
        const int nc = s_nc;
        n2_forward(T, L, N, nc, L.chg);
        if (tid == 0) {
            double tot = 0.0;
            int q = 0;
            for (int c = 0; c < N; ++c) {
                if (q < nc && L.chg[q] == c) {
                    tot += L.eatn[c];
                    ++q;
                } else {
                    tot += L.eat[c];
                }
            }
            s_en = tot;
        }
        __syncthreads();
        // Metropolis on the energy change (src/Exchange.jl:53-57, :74-83)
        const double new_en = s_en;
        const double dd = new_en - E;
        const bool accept = (dd <= 0.0) || (det_exp(-dd * beta) >= d0.umet);
        if (tid == 0 && a.trace_de != nullptr && a.trace_cycle < a.trace_cycles) {
            const size_t ti = ((size_t)Tg * a.trace_cycles + a.trace_cycle) * nsteps + m;
            a.trace_de[ti] = dd;
            a.trace_acc[ti] = (int8_t)accept;
            a.trace_idx[ti] = swap ? N + ma[0] * N + ma[1] : ma[0];
        }
        if (swap) nsw += 1;
        if (accept) {
            // swap_atom_config! / swap_move_config! with swap_vars!(NNPVariables2a)
            for (int e = tid; e < 3 * N; e += NNP2_NT) L.X[e] = L.XT[e];
            for (int e = tid; e < nm * N; e += NNP2_NT) {
                const int mm = e / N, k = e % N, aa = ma[mm];
                L.R2[aa * N + k] = L.RN[e];
                L.R2[k * N + aa] = L.RN[e];
                L.FC[aa * N + k] = L.FN[e];
                L.FC[k * N + aa] = L.FN[e];
            }
            for (int e = tid; e < nc * ng; e += NNP2_NT) {
                const int c = L.chg[e / ng], r = e % ng;
                gG[(size_t)c * ng + r] = L.D[(size_t)c * ng + r];
            }
            for (int e = tid; e < nc; e += NNP2_NT) L.eat[L.chg[e]] = L.eatn[L.chg[e]];
            E = new_en;
            if (swap) nsw_acc += 1;
            else nacc += 1;
        }
        if (tid == 0) {
            L.mi[ma[0]] = -1;
            if (swap) L.mi[ma[1]] = -1;
        }
        __syncthreads();
    }

    // periodic full rebuild bounds the drift of the incremental features
    if (a.recompute_every > 0 && ((long long)sweep + 1) % a.recompute_every == 0) {
        __shared__ double s_tot;
        const double Er = n2_rebuild(T, L, N, &s_tot);
        const double dr = __builtin_fabs(E - Er);
        if (tid == 0 && dr > a.slot_drift[slot]) a.slot_drift[slot] = dr;
        for (int e = tid; e < N * ng; e += NNP2_NT) gG[e] = L.D[e];
        for (int e = tid; e < N; e += NNP2_NT) L.eat[e] = L.eatn[e];
        E = Er;
        __syncthreads();
    }

    N2R *gX = reinterpret_cast<N2R *>(a.coords) + (size_t)slot * 3 * NNP_NP;
    for (int e = tid; e < 3 * N; e += NNP2_NT) gX[(e % 3) * NNP_NP + e / 3] = L.X[e];
    for (int e = tid; e < N; e += NNP2_NT) reinterpret_cast<N2R *>(T.eatom)[(size_t)slot * NNP_NP + e] = L.eat[e];
    if (tid == 0) {
        a.slot_en[slot] = E;
        a.t_count_atom[2 * Tg] += nacc;
        a.t_count_atom[2 * Tg + 1] += nacc;
        a.t_count_vol[2 * Tg] += nsw_acc;
        a.t_count_vol[2 * Tg + 1] += nsw;
        if (s_err) a.run_error[a.slot_run[slot]] = 1;
    }
}

}  // namespace

using SweepFn = void (*)(SweepArgs);

SweepFn N2_SWEEP_FN() { return &nnp2_sweep_kernel; }
SweepFn N2_INIT_FN() { return &nnp2_init_kernel; }

}  // namespace ptmc
