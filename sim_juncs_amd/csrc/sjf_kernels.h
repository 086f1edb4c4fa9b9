// sjf_kernels.h -- gfx950 kernels of the FDTD engine (device code only).
//
// Hot path: one fused launch per time step (k_step) that performs, for every cell,
//   H^{n+1/2} = H^{n-1/2} - (dt/dx) (dE + psi_H)                       (Yee curl, CPML)
//   Q_k^{n+1} = alpha_k Q_k^n + beta_k P_k^n + c3_k E^n, P_k^{n+1} = P_k^n + Q_k^{n+1}
//   E^{n+1}   = E^n + inv_eps ((dt/dx) (dH + psi_E) - sum_k Q_k^{n+1})
// This is the E-direct form of meep's D/E split (src/disp.cpp:740 fields.step(): B, H, D,
// polarisation, E): with D = eps_inf E + sum P, D^{n+1} - D^n = dt curl H, and
// Q = P^{n+1} - P^n the textbook Lorentz/Drude recurrence becomes the (P, Q) pair above,
// which needs no D array and, for Drude poles, no P array.  In FP32 a Lorentz pole keeps P^n
// and P^{n-1} (Q^n = P^n - P^{n-1}) in two slots whose roles alternate with the step parity,
// so each step reads both and writes only P^{n+1} over P^{n-1}: 3 values moved per pole and
// component instead of 4 for a stored (P, Q).  FP64 keeps (P, Q): its tile is held to 3
// wavefronts per SIMD by its registers and runs faster without the role selects.  A Drude pole
// keeps Q^n alone (2 values).
//
// Layout in HBM: each component is a dense nx*ny*nz array, z fastest.  E and H and the H
// auxiliaries are double-buffered (read buffer 0, write buffer 1, swapped every step) because
// neighbouring workgroups recompute one halo row of H from the old state; polarisation state
// and the E auxiliaries are only touched by the owning thread and are updated in place.
//
// Work decomposition (MI355X: 64-wide wavefronts, 8 XCDs with private L2):
//   * one workgroup spans the whole z line (blockDim = nz rounded up to 64, <= 1024), so the
//     z neighbours of H come from LDS and the z halo disappears;
//   * each thread owns TY consecutive y rows (register blocking) and recomputes H on one
//     extra row below them (the y halo);
//   * the workgroup marches along x over a chunk of planes, carrying E^n of the next plane
//     and H^{n+1/2} of the previous plane in registers (no x halo except one plane per chunk);
//   * workgroup ids are remapped so that consecutive tiles land on the same XCD and share
//     halo rows in that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sjf {

constexpr int kMaxPoles = 4;
// Lorentz poles stored as (P^n, P^{n-1}) with alternating roles (FP32) or as (P, Q) (FP64)
template <typename T>
constexpr bool kPolePair = sizeof(T) == 4;
constexpr int kDesignMat = 255;
constexpr int kMaxDftFreqs = 64;

template <typename T>
struct MatLut {
    T inv_eps;
    T eps_inf;
    int np;
    int lorentz_mask;
    T alpha[kMaxPoles];
    T beta[kMaxPoles];
    T c3[kMaxPoles];
};

// Kernel arguments of the fused step.  Few base pointers on purpose: the component, pole-slot
// and PML-segment arrays are addressed as offsets from them, so the whole argument block stays
// in scalar registers (a per-array pointer table overflows the SGPR file and spills into
// vector-register lanes).
template <typename T>
struct StepArgs {
    int nx, ny, nz;
    int px, py, pz;
    int nty, xc;
    int rt;       // row-of-threads form: thread rows per workgroup (owned rows + 1 halo row)
    int zseg;     // register form: z cells per pass of a workgroup (nz: one pass)
    long N;       // cells
    T cdt;        // dt / dx
    T* F[2];              // field sets [x][Ex Ey Ez Hx Hy Hz][y][z]; step n reads F[n&1], writes F[~n&1]
    const long* d_step;   // the engine's step counter
    int uniform;  // 1: single non-dispersive material, no material word read
    int nmat;
    const uint32_t* matw;
    const MatLut<T>* lut;
    T* Q;     // [x][slot * 3 + comp][y][z]
    T* P;     // [x][slot * 3 + comp][y][z] (Lorentz slots)
    int nqp;  // polarisation planes per x plane (3 * slots)
    int pml_on[3], plo[3], phi[3];
    long S[3];      // PML slab sizes (cells) per axis
    T* psi[3];      // per axis [6][S]: E aux (2 comps), H aux pair A (2), H aux pair B (2);
                    // step n reads pair A and writes pair B when n is even, the reverse when odd
    const T* coef;  // PML tables: for axis d at coff[d]: bE[n], aE[n], bH[n], aH[n]
    int coff[3];
    int dlo[3], dhi[3];
    const T* dinv;
    const T* emap;  // per-cell 1/eps_inf [x][Ex Ey Ez][y][z] (sjf_set_eps_maps), or null: the LUT's
};

#define GLOBAL __attribute__((address_space(1)))

__device__ __forceinline__ int wrap_or_ghost(int c, int n, int periodic) {
    if (c >= 0 && c < n) return c;
    if (!periodic) return -1;
    // any number of periods away: a tile's rows (and the planes ahead of its march) can lie
    // more than one period past the edge of an axis shorter than the tile
    const int m = c % n;
    return m < 0 ? m + n : m;
}

template <typename T>
__device__ __forceinline__ bool in_pml(const StepArgs<T>& a, int d, int c, int n) {
    return a.pml_on[d] && (c < a.plo[d] || c >= n - a.phi[d]);
}

// CPML update of one auxiliary: psi <- b psi + a d; returns d + psi.
template <typename T>
__device__ __forceinline__ T cpml(T d, T b, T c, T& psi) {
    psi = b * psi + c * d;
    return d + psi;
}

// LDS-scoped workgroup barrier: orders only the LDS exchange, so global loads issued before it
// (next-plane prefetch) and global stores stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---- raw buffer access ----------------------------------------------------------------------
// A raw buffer descriptor over p (no stride, the gfx9 dword data format) whose range is the first
// 2 GiB from p.  A load or store through it names the descriptor (four scalar registers), a 32-bit
// per-lane offset (one vector register) and a 32-bit wave-uniform offset (one scalar register), so
// an access needs no vector address arithmetic and a loop-invariant lane offset stays a single
// register.  A lane whose offset is kOOB lies outside the range: its load returns 0 and its store
// is dropped without a memory access, which predicates single lanes without branching.  Callers
// keep every in-range access (lane plus uniform offset) below 2 GiB from the descriptor base.
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kOOB = 0x80000000u;
__device__ __forceinline__ Rsrc rsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)0x80000000u, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(Rsrc r, uint32_t v, uint32_t s);
template <>
__device__ __forceinline__ float bld<float>(Rsrc r, uint32_t v, uint32_t s) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)s, 0));
}
template <>
__device__ __forceinline__ double bld<double>(Rsrc r, uint32_t v, uint32_t s) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)v, (int)s, 0));
}
template <>
__device__ __forceinline__ uint32_t bld<uint32_t>(Rsrc r, uint32_t v, uint32_t s) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)v, (int)s, 0);
}
// NT: the non-temporal cache policy (a hint that the line will not be read again soon).  A step
// over more data than the caches hold cannot find its writes there at the next step, and the
// hint keeps the written lines from displacing the y-halo rows that neighbouring workgroups
// read again.
template <bool NT = false>
__device__ __forceinline__ void bst(Rsrc r, uint32_t v, uint32_t s, float x) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), r, (int)v, (int)s, NT ? 2 : 0);
}
template <bool NT = false>
__device__ __forceinline__ void bst(Rsrc r, uint32_t v, uint32_t s, double x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, (int)v, (int)s, NT ? 2 : 0);
}
// true in every lane of the wavefront when p holds in at least one active lane (wave-uniform)
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// Field and polarisation layout: x planes of components, [x][comp][y][z] for a field set (six
// components) and [x][slot * 3 + comp][y][z] for Q and P.  One x plane of a set is contiguous,
// so a kernel addresses a plane through one descriptor and a component by a 32-bit offset.
__host__ __device__ __forceinline__ long field_index(long q, int c, long nyz) { return q + (5 * (q / nyz) + c) * nyz; }

// Fused H / polarisation / E update of one time step.  TY: y rows per thread; NP: pole slots
// compiled in (>= the largest pole count of any material).
// A workgroup covers one z line (thread k = cell k) of TY consecutive y rows and marches along
// x.  Each plane iteration issues every global load it needs first (fields, PML auxiliaries,
// polarisation state, and E and material words of the planes ahead), then computes H, exchanges
// the z neighbours of H through LDS, and computes E.
// Addressing: a thread's byte offset within a component plane depends only on its row and z cell,
// so it is fixed for the whole march; each access is a per-plane descriptor, that lane offset and
// a uniform component offset.  Ghost rows, lanes past the end of the z line and lanes outside a
// z PML layer, a pole slot or the design region carry the out-of-range offset instead of
// branching; whole-wavefront conditions (row, plane, x/y PML, any lane in a z layer / using a
// slot / in the design region) are uniform branches.
template <typename T, int TY, int NP, bool ZP, bool NTS>
__device__ __forceinline__ void step_tile(const StepArgs<T>& a, const int t) {
    constexpr uint32_t TS = sizeof(T);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int B = blockDim.x;
    T* sh = reinterpret_cast<T*>(smem_raw);  // [2 parity][2 comp][TY][B]
    MatLut<T>* slut = reinterpret_cast<MatLut<T>*>(sh + 4 * TY * B);
    {
        GLOBAL const uint32_t* src = (GLOBAL const uint32_t*)a.lut;
        uint32_t* dst = reinterpret_cast<uint32_t*>(slut);
        const int words = (int)(sizeof(MatLut<T>) / 4) * a.nmat;
        for (int q = threadIdx.x; q < words; q += B) dst[q] = src[q];
    }

    const long step = *(GLOBAL const long*)a.d_step;
    const char* F0 = (const char*)a.F[step & 1];
    const char* F1 = (const char*)a.F[(step + 1) & 1];
    const char* matw = (const char*)a.matw;
    GLOBAL const T* coef = (GLOBAL const T*)a.coef;
    const uint32_t hr = (step & 1) ? 4 : 2;  // H auxiliary pair read this step
    const uint32_t hw = (step & 1) ? 2 : 4;  // and written
    // FP32 Lorentz pole slots hold P^n and P^{n-1}: P^n in the Q array on even steps and in
    // the P array on odd ones, P^{n+1} overwrites P^{n-1}.  Drude slots hold Q^n in the Q array.
    const bool podd = (step & 1) != 0;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int tyi = t % a.nty, xci = t / a.nty;
    const int j0 = tyi * TY;
    const int x0 = xci * a.xc;
    const int x1 = min(x0 + a.xc, nx);
    // The z line in passes of zseg cells (one pass when the workgroup spans the whole line).  The
    // last lane of a pass keeps the H it forms in zh, where the first lane of the next pass reads
    // its z neighbour; passes alternate between two halves of zh.
    const int Z = ZP ? a.zseg : nz;  // ZP: compiled for z passes
    const int npass = ZP ? (nz + Z - 1) / Z : 1;
    T* zh = reinterpret_cast<T*>(slut + a.nmat);  // [2 pass parity][xc][TY][Hx, Hy]
    const int kl = threadIdx.x;
    for (int pass = 0; pass < npass; ++pass) {
        const int k = pass * Z + kl;
        const bool kin = kl < Z && k < nz;
        const int kp1 = wrap_or_ghost(k + 1, nz, a.pz);
        const uint32_t PB = (uint32_t)ny * (uint32_t)nz * TS;  // bytes of one component plane
        const long FPB = 6L * PB;                             // bytes of one x plane of a field set
        const long QPB = (long)a.nqp * PB;                    // and of the polarisation
        const long PW = (long)ny * nz * 4;                    // and of the material words
        const T cdt = a.cdt;
        const bool uniform = a.uniform != 0;
        const int Ly = a.plo[1] + a.phi[1], Lz = a.plo[2] + a.phi[2];
        const uint32_t SB0 = (uint32_t)(a.S[0] * TS), SB1 = (uint32_t)(a.S[1] * TS), SB2 = (uint32_t)(a.S[2] * TS);
        const Rsrc rpx = rsrc(a.psi[0]), rpy = rsrc(a.psi[1]), rpz = rsrc(a.psi[2]);

        int gj[TY + 2];        // global row of local row r-1 (r = 0 is the halo row below the tile)
        bool rv[TY + 2];       // row inside the grid (uniform)
        uint32_t oR[TY + 2];   // byte offset of (row, k) within a component plane, or kOOB
        uint32_t oW[TY + 2];   // the same for the 4-byte material words
#pragma unroll
        for (int r = 0; r < TY + 2; ++r) {
            gj[r] = wrap_or_ghost(j0 + r - 1, ny, a.py);
            rv[r] = gj[r] >= 0;
            const bool ok = rv[r] && kin;
            oR[r] = ok ? (uint32_t)(gj[r] * nz + k) * TS : kOOB;
            oW[r] = ok ? (uint32_t)(gj[r] * nz + k) * 4u : kOOB;
        }
        uint32_t oK1[TY + 1];  // byte offset of (row, k + 1), or kOOB
#pragma unroll
        for (int r = 0; r < TY + 1; ++r)
            oK1[r] = (rv[r] && kin && kp1 >= 0) ? (uint32_t)(gj[r] * nz + kp1) * TS : kOOB;
        bool own_row[TY];
#pragma unroll
        for (int r = 0; r < TY; ++r) own_row[r] = (j0 + r) < ny;
        bool ypml[TY + 1];
        int sy[TY + 1];
#pragma unroll
        for (int r = 0; r < TY + 1; ++r) {
            ypml[r] = rv[r] && in_pml(a, 1, gj[r], ny);
            sy[r] = ypml[r] ? (gj[r] < a.plo[1] ? gj[r] : gj[r] - (ny - a.phi[1]) + a.plo[1]) : 0;
        }

        const bool zpml = kin && in_pml(a, 2, k, nz);
        const bool zany = wave_any(zpml);  // some lane of this wavefront lies in a z PML layer
        const uint32_t oK = kin ? (uint32_t)k * TS : kOOB;
        const uint32_t oZ = zpml ? (uint32_t)(k < a.plo[2] ? k : k - (nz - a.phi[2]) + a.plo[2]) * TS : kOOB;
        GLOBAL const T* cz = coef + a.coff[2];
        const T bEz = zpml ? cz[k] : T(0), aEz = zpml ? cz[nz + k] : T(0);
        const T bHz = zpml ? cz[2 * nz + k] : T(0), aHz = zpml ? cz[3 * nz + k] : T(0);

        lds_barrier();

        T ec[TY + 2][3];  // E^n on plane i, rows -1..TY
        T en[TY + 2][3];  // E^n on plane i+1
        T hp[TY][2];      // H^{n+1/2} (Hy, Hz) on plane i-1, rows 0..TY-1
        uint32_t mw[TY];  // material words of plane i, owned rows

        const int istart = (x0 > 0 || a.px) ? x0 - 1 : x0;
        {
            const int p0 = wrap_or_ghost(istart, nx, a.px);
            const int p1 = wrap_or_ghost(istart + 1, nx, a.px);
            const Rsrc r0 = rsrc(F0 + (long)(p0 >= 0 ? p0 : 0) * FPB);
            const Rsrc r1 = rsrc(F0 + (long)(p1 >= 0 ? p1 : 0) * FPB);
#pragma unroll
            for (int r = 0; r < TY + 2; ++r) {
                const uint32_t o0 = p0 >= 0 ? oR[r] : kOOB, o1 = p1 >= 0 ? oR[r] : kOOB;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ec[r][c] = bld<T>(r0, o0, c * PB);
                    en[r][c] = bld<T>(r1, o1, c * PB);
                }
            }
            const bool mw0 = !uniform && istart >= x0;
            const Rsrc rm = rsrc(matw + (long)(p0 >= 0 ? p0 : 0) * PW);
#pragma unroll
            for (int r = 0; r < TY; ++r) {
                hp[r][0] = hp[r][1] = T(0);
                mw[r] = 0u;
                if (mw0 && own_row[r]) mw[r] = bld<uint32_t>(rm, oW[r + 1], 0);
            }
        }

        for (int i = istart; i < x1; ++i) {
            const bool own_plane = i >= x0;
            const bool has_next = i + 1 < x1;
            const int pi = wrap_or_ghost(i, nx, a.px);
            const int pn = wrap_or_ghost(i + 1, nx, a.px);
            const int pnn = wrap_or_ghost(i + 2, nx, a.px);
            const bool xpml = in_pml(a, 0, pi, nx);
            const uint32_t sxo = xpml ? (uint32_t)(pi < a.plo[0] ? pi : pi - (nx - a.phi[0]) + a.plo[0]) * PB : 0u;
            GLOBAL const T* cx = coef + a.coff[0];
            const T bEx = xpml ? cx[pi] : T(0), aEx = xpml ? cx[nx + pi] : T(0);
            const T bHx = xpml ? cx[2 * nx + pi] : T(0), aHx = xpml ? cx[3 * nx + pi] : T(0);
            const Rsrc rF = rsrc(F0 + (long)pi * FPB);
            const Rsrc rW = rsrc(F1 + (long)pi * FPB);

            // ---------------------------------------------------------------- loads ----------
            T enn[TY + 2][3];
            uint32_t mwn[TY];
            {
                // plane i + 2 (a ghost plane reads as zero; after the last plane nothing is read)
                const bool nn_ok = has_next && pnn >= 0;
                const Rsrc rn = rsrc(F0 + (long)(nn_ok ? pnn : pi) * FPB);
#pragma unroll
                for (int r = 0; r < TY + 2; ++r) {
                    const uint32_t o = nn_ok ? oR[r] : kOOB;
                    enn[r][0] = bld<T>(rn, o, 0);
                    enn[r][1] = (r < TY + 1) ? bld<T>(rn, o, PB) : T(0);
                    enn[r][2] = bld<T>(rn, o, 2 * PB);
                }
                const bool mwn_ok = !uniform && has_next && pn >= 0;
                const Rsrc rm = rsrc(matw + (long)(mwn_ok ? pn : pi) * PW);
#pragma unroll
                for (int r = 0; r < TY; ++r) {
                    mwn[r] = 0u;
                    if (mwn_ok && own_row[r]) mwn[r] = bld<uint32_t>(rm, oW[r + 1], 0);
                }
            }
            T h0[TY + 1][3], ek1[TY + 1][2], pH[TY + 1][3][2];
#pragma unroll
            for (int r = 0; r < TY + 1; ++r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) h0[r][c] = bld<T>(rF, oR[r], (3 + c) * PB);
#pragma unroll
                for (int c = 0; c < 2; ++c) ek1[r][c] = bld<T>(rF, oK1[r], c * PB);
#pragma unroll
                for (int d = 0; d < 3; ++d) pH[r][d][0] = pH[r][d][1] = T(0);
                if (rv[r] && xpml) {
                    pH[r][0][0] = bld<T>(rpx, oR[r], hr * SB0 + sxo);
                    pH[r][0][1] = bld<T>(rpx, oR[r], (hr + 1) * SB0 + sxo);
                }
                if (ypml[r]) {
                    const uint32_t so = (uint32_t)((pi * Ly + sy[r]) * nz) * TS;
                    pH[r][1][0] = bld<T>(rpy, oK, hr * SB1 + so);
                    pH[r][1][1] = bld<T>(rpy, oK, (hr + 1) * SB1 + so);
                }
                if (rv[r] && zany) {
                    const uint32_t so = (uint32_t)((pi * ny + gj[r]) * Lz) * TS;
                    pH[r][2][0] = bld<T>(rpz, oZ, hr * SB2 + so);
                    pH[r][2][1] = bld<T>(rpz, oZ, (hr + 1) * SB2 + so);
                }
            }
            T pE[TY][3][2];
            T Qv[TY][3][NP > 0 ? NP : 1], Pv[TY][3][NP > 0 ? NP : 1];
            const Rsrc rQ = rsrc((const char*)a.Q + (long)pi * QPB);
            const Rsrc rP = rsrc((const char*)a.P + (long)pi * QPB);
#pragma unroll
            for (int r = 0; r < TY; ++r) {
#pragma unroll
                for (int d = 0; d < 3; ++d) pE[r][d][0] = pE[r][d][1] = T(0);
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int s = 0; s < (NP > 0 ? NP : 1); ++s) Qv[r][c][s] = Pv[r][c][s] = T(0);
                if (!(own_plane && own_row[r])) continue;
                const uint32_t o = oR[r + 1];
                if (xpml) {
                    pE[r][0][0] = bld<T>(rpx, o, sxo);
                    pE[r][0][1] = bld<T>(rpx, o, SB0 + sxo);
                }
                if (ypml[r + 1]) {
                    const uint32_t so = (uint32_t)((pi * Ly + sy[r + 1]) * nz) * TS;
                    pE[r][1][0] = bld<T>(rpy, oK, so);
                    pE[r][1][1] = bld<T>(rpy, oK, SB1 + so);
                }
                if (zany) {
                    const uint32_t so = (uint32_t)((pi * ny + gj[r + 1]) * Lz) * TS;
                    pE[r][2][0] = bld<T>(rpz, oZ, so);
                    pE[r][2][1] = bld<T>(rpz, oZ, SB2 + so);
                }
                if (NP > 0 && !uniform) {
                    int np[3], lm[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const int m = (mw[r] >> (8 * c)) & 0xff;
                        const int mc = m == kDesignMat ? 0 : m;  // LUT index (design voxels have no poles)
                        np[c] = m == kDesignMat ? 0 : slut[mc].np;
                        lm[c] = slut[mc].lorentz_mask;
                    }
#pragma unroll
                    for (int s = 0; s < NP; ++s) {
                        if (!wave_any(s < np[0] || s < np[1] || s < np[2])) continue;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            Qv[r][c][s] = bld<T>(rQ, s < np[c] ? o : kOOB, (s * 3 + c) * PB);
                            Pv[r][c][s] = bld<T>(rP, (s < np[c] && ((lm[c] >> s) & 1)) ? o : kOOB, (s * 3 + c) * PB);
                        }
                    }
                }
            }

            // ---------------------------------------------------------------- H^{n+1/2} -------
            T h[TY + 1][3];
#pragma unroll
            for (int r = 0; r < TY + 1; ++r) {
                const bool ok = rv[r];
                const int j = gj[r];
                const bool wr = own_plane && r >= 1 && own_row[r > 0 ? r - 1 : 0] && ok;
                T dEz_y = ec[r + 1][2] - ec[r][2];
                T dEx_y = ec[r + 1][0] - ec[r][0];
                T dEy_z = ek1[r][1] - ec[r][1];
                T dEx_z = ek1[r][0] - ec[r][0];
                T dEz_x = en[r][2] - ec[r][2];
                T dEy_x = en[r][1] - ec[r][1];
                if (ypml[r]) {
                    GLOBAL const T* cy = coef + a.coff[1];
                    const T b = cy[2 * ny + j], c = cy[3 * ny + j];
                    dEz_y = cpml(dEz_y, b, c, pH[r][1][0]);
                    dEx_y = cpml(dEx_y, b, c, pH[r][1][1]);
                    if (wr) {
                        const uint32_t so = (uint32_t)((pi * Ly + sy[r]) * nz) * TS;
                        bst<NTS>(rpy, oK, hw * SB1 + so, pH[r][1][0]);
                        bst<NTS>(rpy, oK, (hw + 1) * SB1 + so, pH[r][1][1]);
                    }
                }
                if (ok && zany) {  // lanes outside the layers: b = a = 0 leaves the term unchanged
                    dEy_z = cpml(dEy_z, bHz, aHz, pH[r][2][0]);
                    dEx_z = cpml(dEx_z, bHz, aHz, pH[r][2][1]);
                    if (wr) {
                        const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
                        bst<NTS>(rpz, oZ, hw * SB2 + so, pH[r][2][0]);
                        bst<NTS>(rpz, oZ, (hw + 1) * SB2 + so, pH[r][2][1]);
                    }
                }
                if (ok && xpml) {
                    dEz_x = cpml(dEz_x, bHx, aHx, pH[r][0][0]);
                    dEy_x = cpml(dEy_x, bHx, aHx, pH[r][0][1]);
                    if (wr) {
                        bst<NTS>(rpx, oR[r], hw * SB0 + sxo, pH[r][0][0]);
                        bst<NTS>(rpx, oR[r], (hw + 1) * SB0 + sxo, pH[r][0][1]);
                    }
                }
                // outside a metallic edge H is identically zero (ghost row of the halo)
                h[r][0] = ok ? h0[r][0] - cdt * (dEz_y - dEy_z) : T(0);
                h[r][1] = ok ? h0[r][1] - cdt * (dEx_z - dEz_x) : T(0);
                h[r][2] = ok ? h0[r][2] - cdt * (dEy_x - dEx_y) : T(0);
                if (wr) {
                    bst<NTS>(rW, oR[r], 3 * PB, h[r][0]);
                    bst<NTS>(rW, oR[r], 4 * PB, h[r][1]);
                    bst<NTS>(rW, oR[r], 5 * PB, h[r][2]);
                }
            }

            if (own_plane) {
                T* buf = sh + (i & 1) * 2 * TY * B;
#pragma unroll
                for (int r = 0; r < TY; ++r) {
                    buf[r * B + kl] = h[r + 1][0];
                    buf[(TY + r) * B + kl] = h[r + 1][1];
                }
                if (ZP && pass + 1 < npass && kl == Z - 1) {
                    T* zo = zh + ((pass & 1) * a.xc + (i - x0)) * 2 * TY;
#pragma unroll
                    for (int r = 0; r < TY; ++r) {
                        zo[2 * r] = h[r + 1][0];
                        zo[2 * r + 1] = h[r + 1][1];
                    }
                }
                lds_barrier();

                // ------------------------------------------------------------ E^{n+1} ---------
#pragma unroll
                for (int r = 0; r < TY; ++r) {
                    if (!own_row[r]) continue;
                    const int j = gj[r + 1];
                    const uint32_t o = oR[r + 1];
                    // z neighbour below: the lane below in this pass, else the previous
                    // pass's last lane (zh), else the periodic image (one pass), else a ghost
                    T hxkm = T(0), hykm = T(0);
                    if (kl > 0) {
                        hxkm = buf[r * B + kl - 1];
                        hykm = buf[(TY + r) * B + kl - 1];
                    } else if (ZP && pass > 0) {
                        const T* zi = zh + (((pass - 1) & 1) * a.xc + (i - x0)) * 2 * TY;
                        hxkm = zi[2 * r];
                        hykm = zi[2 * r + 1];
                    } else if (a.pz) {  // periodic z runs one pass: the image of k = 0 is nz - 1
                        hxkm = buf[r * B + nz - 1];
                        hykm = buf[(TY + r) * B + nz - 1];
                    }
                    T dHz_y = h[r + 1][2] - h[r][2];
                    T dHx_y = h[r + 1][0] - h[r][0];
                    T dHy_z = h[r + 1][1] - hykm;
                    T dHx_z = h[r + 1][0] - hxkm;
                    T dHz_x = h[r + 1][2] - hp[r][1];
                    T dHy_x = h[r + 1][1] - hp[r][0];
                    if (ypml[r + 1]) {
                        GLOBAL const T* cy = coef + a.coff[1];
                        const T b = cy[j], c = cy[ny + j];
                        dHz_y = cpml(dHz_y, b, c, pE[r][1][0]);
                        dHx_y = cpml(dHx_y, b, c, pE[r][1][1]);
                        const uint32_t so = (uint32_t)((pi * Ly + sy[r + 1]) * nz) * TS;
                        bst<NTS>(rpy, oK, so, pE[r][1][0]);
                        bst<NTS>(rpy, oK, SB1 + so, pE[r][1][1]);
                    }
                    if (zany) {
                        dHy_z = cpml(dHy_z, bEz, aEz, pE[r][2][0]);
                        dHx_z = cpml(dHx_z, bEz, aEz, pE[r][2][1]);
                        const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
                        bst<NTS>(rpz, oZ, so, pE[r][2][0]);
                        bst<NTS>(rpz, oZ, SB2 + so, pE[r][2][1]);
                    }
                    if (xpml) {
                        dHz_x = cpml(dHz_x, bEx, aEx, pE[r][0][0]);
                        dHy_x = cpml(dHy_x, bEx, aEx, pE[r][0][1]);
                        bst<NTS>(rpx, o, sxo, pE[r][0][0]);
                        bst<NTS>(rpx, o, SB0 + sxo, pE[r][0][1]);
                    }
                    const T curl[3] = {dHz_y - dHy_z, dHx_z - dHz_x, dHy_x - dHx_y};
                    if (uniform) {
                        const T inv = slut[0].inv_eps;
#pragma unroll
                        for (int c = 0; c < 3; ++c) bst<NTS>(rW, o, c * PB, ec[r + 1][c] + inv * (cdt * curl[c]));
                    } else {
                        int mm[3];
                        bool des[3];
                        T inv[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const int m = (mw[r] >> (8 * c)) & 0xff;
                            des[c] = m == kDesignMat;
                            mm[c] = des[c] ? 0 : m;
                            inv[c] = slut[mm[c]].inv_eps;
                        }
                        if (a.emap) {  // per-cell eps_inf (smoothed maps)
                            const Rsrc re = rsrc((const char*)a.emap + (long)pi * 3 * PB);
#pragma unroll
                            for (int c = 0; c < 3; ++c) inv[c] = bld<T>(re, o, c * PB);
                        }
                        if (wave_any(des[0] || des[1] || des[2])) {
                            // design voxels: 1/eps from the design array [dx][dy][dz]
                            const Rsrc rd = rsrc(a.dinv);
                            const uint32_t dq = (uint32_t)(((pi - a.dlo[0]) * (a.dhi[1] - a.dlo[1]) + (j - a.dlo[1])) *
                                                               (a.dhi[2] - a.dlo[2]) + (kl + pass * Z - a.dlo[2])) * TS;
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const T v = bld<T>(rd, des[c] ? dq : kOOB, 0);
                                inv[c] = des[c] ? v : inv[c];
                            }
                        }
                        T sq[3] = {T(0), T(0), T(0)};
#pragma unroll
                        for (int s = 0; s < NP; ++s) {
                            int npl[3];
#pragma unroll
                            for (int c = 0; c < 3; ++c) npl[c] = des[c] ? 0 : slut[mm[c]].np;
                            if (!wave_any(s < npl[0] || s < npl[1] || s < npl[2])) continue;
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const MatLut<T>& L = slut[mm[c]];
                                const bool vq = s < npl[c];
                                const bool vp = vq && ((L.lorentz_mask >> s) & 1);
                                T qn;
                                if constexpr (kPolePair<T>) {
                                    const T cur = podd ? Pv[r][c][s] : Qv[r][c][s];  // Lorentz: P^n
                                    const T old = podd ? Qv[r][c][s] : Pv[r][c][s];  // and P^{n-1}
                                    qn = vp ? L.alpha[s] * (cur - old) + L.beta[s] * cur + L.c3[s] * ec[r + 1][c]
                                            : L.alpha[s] * Qv[r][c][s] + L.c3[s] * ec[r + 1][c];
                                    qn = vq ? qn : T(0);
                                    bst<NTS>(rP, (vp && !podd) ? o : kOOB, (s * 3 + c) * PB, cur + qn);
                                    bst<NTS>(rQ, (vq && !(vp && !podd)) ? o : kOOB, (s * 3 + c) * PB, vp ? cur + qn : qn);
                                } else {
                                    // Pv is zero unless vp (out-of-range load), so beta * Pv adds nothing then
                                    qn = L.alpha[s] * Qv[r][c][s] + L.c3[s] * ec[r + 1][c] + L.beta[s] * Pv[r][c][s];
                                    qn = vq ? qn : T(0);
                                    bst<NTS>(rP, vp ? o : kOOB, (s * 3 + c) * PB, Pv[r][c][s] + qn);
                                    bst<NTS>(rQ, vq ? o : kOOB, (s * 3 + c) * PB, qn);
                                }
                                sq[c] += qn;
                            }
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) bst<NTS>(rW, o, c * PB, ec[r + 1][c] + inv[c] * (cdt * curl[c] - sq[c]));
                    }
                }
            }

#pragma unroll
            for (int r = 0; r < TY; ++r) {
                hp[r][0] = h[r + 1][1];
                hp[r][1] = h[r + 1][2];
                mw[r] = mwn[r];
            }
#pragma unroll
            for (int r = 0; r < TY + 2; ++r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ec[r][c] = en[r][c];
                    en[r][c] = enn[r][c];
                }
            }
        }
    }  // pass
}

// Fused step, row-of-threads form.  A workgroup holds a.rt rows of nz threads (one cell of a
// y row per thread); thread row 0 recomputes H on the row below the tile (the y halo) and rows
// 1..rt-1 own rows j0..j0+rt-2.  The y and z neighbours of E (for H) and of H (for E) are
// exchanged through LDS, so each field value is loaded from global memory once per tile and
// the y halo costs one extra row per rt-1 owned rows instead of one per row.  The tile
// marches along x; E of the next two planes and H of the next plane are prefetched into
// registers one iteration ahead.  The top thread row loads E of the row above the tile
// (Ex, Ez) itself.  Addressing as in step_tile: per-plane descriptors and 32-bit offsets.
template <typename T, int NP>
__device__ __forceinline__ void step_rows(const StepArgs<T>& a, const int t) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int B = blockDim.x;
    T* sE = reinterpret_cast<T*>(smem_raw);  // [3][B] E^n of the current plane
    T* sH = sE + 3 * B;                       // [3][B] H^{n+1/2} of the current plane
    MatLut<T>* slut = reinterpret_cast<MatLut<T>*>(sH + 3 * B);
    {
        GLOBAL const uint32_t* src = (GLOBAL const uint32_t*)a.lut;
        uint32_t* dst = reinterpret_cast<uint32_t*>(slut);
        const int words = (int)(sizeof(MatLut<T>) / 4) * a.nmat;
        for (int q = threadIdx.x; q < words; q += B) dst[q] = src[q];
    }

    const long step = *(GLOBAL const long*)a.d_step;
    const char* F0 = (const char*)a.F[step & 1];
    const char* F1 = (const char*)a.F[(step + 1) & 1];
    const char* matw = (const char*)a.matw;
    GLOBAL const T* coef = (GLOBAL const T*)a.coef;
    const uint32_t hr = (step & 1) ? 4 : 2;  // H auxiliary pair read this step
    const uint32_t hw = (step & 1) ? 2 : 4;  // and written
    // FP32 Lorentz pole slots hold P^n and P^{n-1}: P^n in the Q array on even steps and in
    // the P array on odd ones, P^{n+1} overwrites P^{n-1}.  Drude slots hold Q^n in the Q array.
    const bool podd = (step & 1) != 0;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int RT = a.rt, R = RT - 1;
    const int tid = threadIdx.x;
    const bool act = tid < RT * nz;
    const int r = act ? tid / nz : 0;
    const int k = act ? tid - r * nz : 0;
    const int tyi = t % a.nty, xci = t / a.nty;
    const int jr = tyi * R + r - 1;  // unwrapped global row of this thread row
    const int j = act ? wrap_or_ghost(jr, ny, a.py) : -1;
    const bool valid = j >= 0;
    const bool owned = act && r >= 1 && jr < ny;  // implies valid
    const bool top = act && r == RT - 1;
    const int jt = top ? wrap_or_ghost(jr + 1, ny, a.py) : -1;
    const int kp1 = wrap_or_ghost(k + 1, nz, a.pz);
    const int km1 = wrap_or_ghost(k - 1, nz, a.pz);
    const T cdt = a.cdt;
    const bool uniform = a.uniform != 0;
    constexpr uint32_t W = sizeof(T);
    const uint32_t PB = (uint32_t)ny * (uint32_t)nz * W;  // bytes of one component plane
    const long FPB = 6L * PB;                            // bytes of one x plane of a field set
    const long QPB = (long)a.nqp * PB;                   // and of the polarisation
    const long PW = (long)ny * nz * 4;                   // and of the material words
    // lane offsets within a component plane; kOOB for lanes that neither read nor write
    const uint32_t oc = valid ? ((uint32_t)j * nz + k) * W : kOOB;
    const uint32_t ot = jt >= 0 ? ((uint32_t)jt * nz + k) * W : kOOB;
    const uint32_t ow = owned ? ((uint32_t)j * nz + k) * 4u : kOOB;
    const uint32_t oo = owned ? oc : kOOB;
    const uint32_t SB0 = (uint32_t)(a.S[0] * W), SB1 = (uint32_t)(a.S[1] * W), SB2 = (uint32_t)(a.S[2] * W);
    const Rsrc rpx = rsrc(a.psi[0]), rpy = rsrc(a.psi[1]), rpz = rsrc(a.psi[2]);
    const int Ly = a.plo[1] + a.phi[1], Lz = a.plo[2] + a.phi[2];

    const bool ypml = valid && in_pml(a, 1, j, ny);
    const bool zpml = valid && in_pml(a, 2, k, nz);
    // wavefront-uniform summaries of the per-lane conditions
    const bool yany = wave_any(ypml), zany = wave_any(zpml), tany = wave_any(jt >= 0), oany = wave_any(owned);
    T bEy = T(0), aEy = T(0), bHy = T(0), aHy = T(0);
    if (ypml) {
        GLOBAL const T* cy = coef + a.coff[1];
        bEy = cy[j];
        aEy = cy[ny + j];
        bHy = cy[2 * ny + j];
        aHy = cy[3 * ny + j];
    }
    T bEz = T(0), aEz = T(0), bHz = T(0), aHz = T(0);
    if (zpml) {
        GLOBAL const T* cz = coef + a.coff[2];
        bEz = cz[k];
        aEz = cz[nz + k];
        bHz = cz[2 * nz + k];
        aHz = cz[3 * nz + k];
    }
    const int syl = ypml ? (j < a.plo[1] ? j : j - (ny - a.phi[1]) + a.plo[1]) : 0;
    const int szl = zpml ? (k < a.plo[2] ? k : k - (nz - a.phi[2]) + a.plo[2]) : 0;
    lds_barrier();  // LUT in LDS: the first plane's pole loads read it before the plane's barriers

    const int x0 = xci * a.xc;
    const int x1 = min(x0 + a.xc, nx);
    const int istart = (x0 > 0 || a.px) ? x0 - 1 : x0;

    T ec[3], en[3], te[2] = {T(0), T(0)}, ten[2] = {T(0), T(0)}, hc[3], hp[2] = {T(0), T(0)};
    uint32_t mw = 0;
    {
        const int p0 = wrap_or_ghost(istart, nx, a.px);
        const int p1 = wrap_or_ghost(istart + 1, nx, a.px);
        const Rsrc r0 = rsrc(F0 + (long)p0 * FPB);
        const Rsrc r1 = rsrc(F0 + (long)(p1 < 0 ? 0 : p1) * FPB);
        const uint32_t o1 = p1 >= 0 ? oc : kOOB;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ec[c] = bld<T>(r0, oc, c * PB);
            en[c] = bld<T>(r1, o1, c * PB);
            hc[c] = bld<T>(r0, oc, (3 + c) * PB);
        }
        // the row above the tile: only lanes of the top thread row load it
        if (tany) {
            const uint32_t t1 = p1 >= 0 ? ot : kOOB;
            te[0] = bld<T>(r0, ot, 0);
            te[1] = bld<T>(r0, ot, 2 * PB);
            ten[0] = bld<T>(r1, t1, 0);
            ten[1] = bld<T>(r1, t1, 2 * PB);
        }
        if (!uniform && istart >= x0 && oany) mw = bld<uint32_t>(rsrc(matw + (long)p0 * PW), ow, 0);
    }

    for (int i = istart; i < x1; ++i) {
        const bool own_plane = i >= x0;
        const bool has_next = i + 1 < x1;
        const int pi = wrap_or_ghost(i, nx, a.px);
        const Rsrc rW = rsrc(F1 + (long)pi * FPB);
        const uint32_t wo = own_plane ? oo : kOOB;  // lanes that write this plane
        const bool wany = own_plane && oany;

        // ------------------------------------------------------------ loads -----------------
        T enn[3] = {T(0), T(0), T(0)}, tenn[2] = {T(0), T(0)}, hn[3] = {T(0), T(0), T(0)};
        uint32_t mwn = 0;
        if (has_next) {
            const int pn = wrap_or_ghost(i + 1, nx, a.px);
            const int pnn = wrap_or_ghost(i + 2, nx, a.px);
            const Rsrc rn = rsrc(F0 + (long)pn * FPB);
            const Rsrc rnn = rsrc(F0 + (long)(pnn < 0 ? 0 : pnn) * FPB);
            const uint32_t onn = pnn >= 0 ? oc : kOOB;
#pragma unroll
            for (int c = 0; c < 3; ++c) enn[c] = bld<T>(rnn, onn, c * PB);
            if (tany) {
                const uint32_t tn = pnn >= 0 ? ot : kOOB;
                tenn[0] = bld<T>(rnn, tn, 0);
                tenn[1] = bld<T>(rnn, tn, 2 * PB);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) hn[c] = bld<T>(rn, oc, (3 + c) * PB);
            if (!uniform && oany) mwn = bld<uint32_t>(rsrc(matw + (long)pn * PW), ow, 0);
        }
        const bool xpml = in_pml(a, 0, pi, nx);
        T bEx = T(0), aEx = T(0), bHx = T(0), aHx = T(0);
        uint32_t sxo = 0;
        if (xpml) {
            GLOBAL const T* cx = coef + a.coff[0];
            bEx = cx[pi];
            aEx = cx[nx + pi];
            bHx = cx[2 * nx + pi];
            aHx = cx[3 * nx + pi];
            sxo = (uint32_t)(pi < a.plo[0] ? pi : pi - (nx - a.phi[0]) + a.plo[0]) * PB;
        }
        // slab lane offsets of this plane (kOOB outside the layer)
        const uint32_t oy = ypml ? (uint32_t)((pi * Ly + syl) * nz + k) * W : kOOB;
        const uint32_t oz = zpml ? (uint32_t)((pi * ny + j) * Lz + szl) * W : kOOB;
        const uint32_t oyw = own_plane && owned ? oy : kOOB, ozw = own_plane && owned ? oz : kOOB;
        T pH[3][2], pE[3][2];
#pragma unroll
        for (int d = 0; d < 3; ++d) pH[d][0] = pH[d][1] = pE[d][0] = pE[d][1] = T(0);
        if (xpml) {
            pH[0][0] = bld<T>(rpx, oc, hr * SB0 + sxo);
            pH[0][1] = bld<T>(rpx, oc, (hr + 1) * SB0 + sxo);
            if (wany) {
                pE[0][0] = bld<T>(rpx, wo, sxo);
                pE[0][1] = bld<T>(rpx, wo, SB0 + sxo);
            }
        }
        if (yany) {
            pH[1][0] = bld<T>(rpy, oy, hr * SB1);
            pH[1][1] = bld<T>(rpy, oy, (hr + 1) * SB1);
            if (wany) {
                pE[1][0] = bld<T>(rpy, oyw, 0);
                pE[1][1] = bld<T>(rpy, oyw, SB1);
            }
        }
        if (zany) {
            pH[2][0] = bld<T>(rpz, oz, hr * SB2);
            pH[2][1] = bld<T>(rpz, oz, (hr + 1) * SB2);
            if (wany) {
                pE[2][0] = bld<T>(rpz, ozw, 0);
                pE[2][1] = bld<T>(rpz, ozw, SB2);
            }
        }
        T Qv[3][NP > 0 ? NP : 1], Pv[3][NP > 0 ? NP : 1];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int s = 0; s < (NP > 0 ? NP : 1); ++s) Qv[c][s] = Pv[c][s] = T(0);
        const Rsrc rQ = rsrc((const char*)a.Q + (long)pi * QPB);
        const Rsrc rP = rsrc((const char*)a.P + (long)pi * QPB);
        int mm[3] = {0, 0, 0}, npl[3] = {0, 0, 0};
        bool des[3] = {false, false, false};
        if (NP > 0 && !uniform && wany) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int m = (mw >> (8 * c)) & 0xff;
                des[c] = m == kDesignMat;
                mm[c] = des[c] ? 0 : m;
                npl[c] = des[c] ? 0 : slut[mm[c]].np;
            }
#pragma unroll
            for (int s = 0; s < NP; ++s) {
                if (!wave_any(s < npl[0] || s < npl[1] || s < npl[2])) continue;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const bool vq = s < npl[c];
                    const bool vp = vq && ((slut[mm[c]].lorentz_mask >> s) & 1);
                    Qv[c][s] = bld<T>(rQ, vq ? wo : kOOB, (s * 3 + c) * PB);
                    Pv[c][s] = bld<T>(rP, vp ? wo : kOOB, (s * 3 + c) * PB);
                }
            }
        }

        // ------------------------------------------------------------ H^{n+1/2} on plane i --
        sE[tid] = ec[0];
        sE[B + tid] = ec[1];
        sE[2 * B + tid] = ec[2];
        lds_barrier();
        T h[3];
        {
            const int up = (top || !act) ? tid : tid + nz;  // the row above within the tile (top row: te)
            const T exj = top ? te[0] : sE[up];
            const T ezj = top ? te[1] : sE[2 * B + up];
            const T exk = kp1 >= 0 ? sE[r * nz + kp1] : T(0);
            const T eyk = kp1 >= 0 ? sE[B + r * nz + kp1] : T(0);
            T dEz_y = ezj - ec[2];
            T dEx_y = exj - ec[0];
            T dEy_z = eyk - ec[1];
            T dEx_z = exk - ec[0];
            T dEz_x = en[2] - ec[2];
            T dEy_x = en[1] - ec[1];
            // lanes outside a PML layer carry zero coefficients and auxiliaries: no change
            if (yany) {
                dEz_y = cpml(dEz_y, bHy, aHy, pH[1][0]);
                dEx_y = cpml(dEx_y, bHy, aHy, pH[1][1]);
            }
            if (zany) {
                dEy_z = cpml(dEy_z, bHz, aHz, pH[2][0]);
                dEx_z = cpml(dEx_z, bHz, aHz, pH[2][1]);
            }
            if (xpml) {
                dEz_x = cpml(dEz_x, bHx, aHx, pH[0][0]);
                dEy_x = cpml(dEy_x, bHx, aHx, pH[0][1]);
            }
            // outside a metallic edge H is identically zero (ghost rows)
            h[0] = valid ? hc[0] - cdt * (dEz_y - dEy_z) : T(0);
            h[1] = valid ? hc[1] - cdt * (dEx_z - dEz_x) : T(0);
            h[2] = valid ? hc[2] - cdt * (dEy_x - dEx_y) : T(0);
            if (wany) {
                bst(rW, wo, 3 * PB, h[0]);
                bst(rW, wo, 4 * PB, h[1]);
                bst(rW, wo, 5 * PB, h[2]);
                if (xpml) {
                    bst(rpx, wo, hw * SB0 + sxo, pH[0][0]);
                    bst(rpx, wo, (hw + 1) * SB0 + sxo, pH[0][1]);
                }
                if (yany) {
                    bst(rpy, oyw, hw * SB1, pH[1][0]);
                    bst(rpy, oyw, (hw + 1) * SB1, pH[1][1]);
                }
                if (zany) {
                    bst(rpz, ozw, hw * SB2, pH[2][0]);
                    bst(rpz, ozw, (hw + 1) * SB2, pH[2][1]);
                }
            }
        }
        sH[tid] = h[0];
        sH[B + tid] = h[1];
        sH[2 * B + tid] = h[2];
        lds_barrier();

        // ------------------------------------------------------------ E^{n+1} on plane i ----
        if (wany) {
            const int dn = owned ? tid - nz : tid;  // the row below within the tile
            const T hxj = sH[dn];
            const T hzj = sH[2 * B + dn];
            const T hxk = km1 >= 0 ? sH[r * nz + km1] : T(0);
            const T hyk = km1 >= 0 ? sH[B + r * nz + km1] : T(0);
            T dHz_y = h[2] - hzj;
            T dHx_y = h[0] - hxj;
            T dHy_z = h[1] - hyk;
            T dHx_z = h[0] - hxk;
            T dHz_x = h[2] - hp[1];
            T dHy_x = h[1] - hp[0];
            if (yany) {
                dHz_y = cpml(dHz_y, bEy, aEy, pE[1][0]);
                dHx_y = cpml(dHx_y, bEy, aEy, pE[1][1]);
                bst(rpy, oyw, 0, pE[1][0]);
                bst(rpy, oyw, SB1, pE[1][1]);
            }
            if (zany) {
                dHy_z = cpml(dHy_z, bEz, aEz, pE[2][0]);
                dHx_z = cpml(dHx_z, bEz, aEz, pE[2][1]);
                bst(rpz, ozw, 0, pE[2][0]);
                bst(rpz, ozw, SB2, pE[2][1]);
            }
            if (xpml) {
                dHz_x = cpml(dHz_x, bEx, aEx, pE[0][0]);
                dHy_x = cpml(dHy_x, bEx, aEx, pE[0][1]);
                bst(rpx, wo, sxo, pE[0][0]);
                bst(rpx, wo, SB0 + sxo, pE[0][1]);
            }
            const T curl[3] = {dHz_y - dHy_z, dHx_z - dHz_x, dHy_x - dHx_y};
            if (uniform) {
                const T inv = slut[0].inv_eps;
#pragma unroll
                for (int c = 0; c < 3; ++c) bst(rW, wo, c * PB, ec[c] + inv * (cdt * curl[c]));
            } else {
                T inv[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    if (NP == 0) {  // material ids were not decoded for the pole loads
                        const int m = (mw >> (8 * c)) & 0xff;
                        des[c] = m == kDesignMat;
                        mm[c] = des[c] ? 0 : m;
                    }
                    inv[c] = slut[mm[c]].inv_eps;
                }
                if (a.emap) {  // per-cell eps_inf (smoothed maps)
                    const Rsrc re = rsrc((const char*)a.emap + (long)pi * 3 * PB);
#pragma unroll
                    for (int c = 0; c < 3; ++c) inv[c] = bld<T>(re, wo, c * PB);
                }
                if (wave_any(des[0] || des[1] || des[2])) {
                    const Rsrc rd = rsrc(a.dinv);
                    const uint32_t dq = (uint32_t)(((pi - a.dlo[0]) * (a.dhi[1] - a.dlo[1]) + (j - a.dlo[1])) *
                                                       (a.dhi[2] - a.dlo[2]) + (k - a.dlo[2])) * W;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T v = bld<T>(rd, des[c] ? dq : kOOB, 0);
                        inv[c] = des[c] ? v : inv[c];
                    }
                }
                T sq[3] = {T(0), T(0), T(0)};
#pragma unroll
                for (int s = 0; s < NP; ++s) {
                    if (!wave_any(s < npl[0] || s < npl[1] || s < npl[2])) continue;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const MatLut<T>& L = slut[mm[c]];
                        const bool vq = s < npl[c];
                        const bool vp = vq && ((L.lorentz_mask >> s) & 1);
                        T qn;
                        if constexpr (kPolePair<T>) {
                            const T cur = podd ? Pv[c][s] : Qv[c][s];  // Lorentz: P^n
                            const T old = podd ? Qv[c][s] : Pv[c][s];  // and P^{n-1}
                            qn = vp ? L.alpha[s] * (cur - old) + L.beta[s] * cur + L.c3[s] * ec[c]
                                    : L.alpha[s] * Qv[c][s] + L.c3[s] * ec[c];
                            qn = vq ? qn : T(0);
                            bst(rP, (vp && !podd) ? wo : kOOB, (s * 3 + c) * PB, cur + qn);
                            bst(rQ, (vq && !(vp && !podd)) ? wo : kOOB, (s * 3 + c) * PB, vp ? cur + qn : qn);
                        } else {
                            // Pv is zero unless vp (out-of-range load), so beta * Pv adds nothing then
                            qn = L.alpha[s] * Qv[c][s] + L.c3[s] * ec[c] + L.beta[s] * Pv[c][s];
                            qn = vq ? qn : T(0);
                            bst(rP, vp ? wo : kOOB, (s * 3 + c) * PB, Pv[c][s] + qn);
                            bst(rQ, vq ? wo : kOOB, (s * 3 + c) * PB, qn);
                        }
                        sq[c] += qn;
                    }
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) bst(rW, wo, c * PB, ec[c] + inv[c] * (cdt * curl[c] - sq[c]));
            }
        }

        hp[0] = h[1];
        hp[1] = h[2];
        mw = mwn;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ec[c] = en[c];
            en[c] = enn[c];
            hc[c] = hn[c];
        }
        te[0] = ten[0];
        te[1] = ten[1];
        ten[0] = tenn[0];
        ten[1] = tenn[1];
    }
}

// XCD-aware tile order: hardware places workgroup w on XCD w % 8; map so that each XCD
// receives a contiguous run of logical tiles (y-tile fastest, then x-chunk), then find the
// instance that owns the tile: instance b owns tile_first[b] <= t < tile_first[b+1].
__device__ __forceinline__ int tile_instance(const int* __restrict__ tile_first, int ninst, int& t) {
    const int w = blockIdx.x, nw = gridDim.x;
    t = (nw % 8 == 0) ? ((w & 7) * (nw >> 3) + (w >> 3)) : w;
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (tile_first[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    t -= tile_first[lo];
    return lo;
}

// Fused step over all instances of a launch set (rows-per-thread form).
// Waves per SIMD the register allocator must leave room for: the FP32 one-row tile with up to
// three pole slots fits 96 registers (5 waves; the z-pass variant keeps 12 bytes per lane in
// scratch); the FP64 one-row tile is held to 168 registers (3 waves; with three pole slots it
// spills about 30 bytes per lane and still runs 3% faster than at 2 waves on the 96^3 box);
// other variants are unconstrained.
template <typename T, int TY, int NP>
constexpr int step_waves() { return (sizeof(T) == 4 && TY == 1 && NP <= 3) ? 5 : (sizeof(T) == 8 && TY == 1) ? 3 : 1; }

// NTS: stores with the non-temporal hint (launch sets that stream more than the caches hold).
template <typename T, int TY, int NP, int MAXT, bool NTS = false>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(step_waves<T, TY, NP>())))
k_step(const StepArgs<T>* __restrict__ args, const int* __restrict__ tile_first, int ninst) {
    int t;
    const int b = tile_instance(tile_first, ninst, t);
    step_tile<T, TY, NP, false, NTS>(args[b], t);
}

// The same with the z line in passes of StepArgs::zseg cells (one row per thread)
template <typename T, int NP, int MAXT, bool NTS = false>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(step_waves<T, 1, NP>())))
k_step_z(const StepArgs<T>* __restrict__ args, const int* __restrict__ tile_first, int ninst) {
    int t;
    const int b = tile_instance(tile_first, ninst, t);
    step_tile<T, 1, NP, true, NTS>(args[b], t);
}

// Fused step over all instances of a launch set (row-of-threads form).
template <typename T, int NP, int MAXT>
__global__ void __launch_bounds__(MAXT) k_step_rows(const StepArgs<T>* __restrict__ args,
                                                    const int* __restrict__ tile_first, int ninst) {
    int t;
    const int b = tile_instance(tile_first, ninst, t);
    step_rows<T, NP>(args[b], t);
}

// ---- per-step scalars live on the device --------------------------------------------------
// Every engine owns a device step counter.  Kernels read the step from it (which also fixes
// the read/write field buffer by its parity) and the last kernel of a step increments it, so a
// sequence of steps can be captured once in a hipGraph and replayed at any step.

__global__ void k_tick(long* const* counters, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i][0] += 1;
}

// The same, folded into the DFT kernel when a step has one (one launch fewer): each workgroup
// signs off after all of its threads have finished (and so have read the counters); the last
// one to sign off advances every counter.  Sign-offs go to eight group words (workgroup index
// mod 8) and the last of each group signs off on a top word, so no word takes more than an
// eighth of the atomics; the words re-arm themselves for the next step.  Each word has a
// 256-byte slot of its own: atomics on one cache line are served one after another, and with
// the eight words in one line a launch of 1300 workgroups spent ~10 us signing off.  No fence
// is needed: no workgroup reads what another writes within the kernel (the counters were read
// before signing off, and the kernel boundary publishes the increment).
// `done` == nullptr: the step ends with k_tick instead.
constexpr int kDoneStride = 64;  // unsigned words per sign-off slot (256 bytes)
constexpr int kDoneWords = 9 * kDoneStride;

struct Tick {
    long* const* counters;
    int n;
    unsigned* done;  // kDoneWords: top word at [0], group g at [(1 + g) * kDoneStride]
};

__device__ __forceinline__ void tick_tail(const Tick& tk) {
    if (!tk.done) return;
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
        const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const unsigned G = nblk < 8u ? nblk : 8u;
        const unsigned g = b % G;
        const unsigned members = (nblk - g + G - 1) / G;  // workgroups of group g
        bool fin = atomicAdd(tk.done + (1 + g) * kDoneStride, 1u) == members - 1;
        if (fin) {
            atomicExch(tk.done + (1 + g) * kDoneStride, 0u);
            fin = atomicAdd(tk.done, 1u) == G - 1;
            if (fin) atomicExch(tk.done, 0u);
        }
        last = fin ? 1 : 0;
    }
    __syncthreads();
    if (last)
        for (int i = threadIdx.x; i < tk.n; i += blockDim.x) tk.counters[i][0] += 1;
}

// ---- source time functions (FP64) ---------------------------------------------------------
// gaussian_src_time_phase::dipole (reference src/disp.cpp:395-400) and meep's
// continuous_src_time::dipole, evaluated in FP64 on the device.
struct PulseDev {
    int kind;  // 0 gaussian, 1 continuous
    int part;  // 0 real part, 1 imaginary part of amp * current
    double omega, width, phi, peak, cutoff;
    double freq, cw_width, t_start, t_end, slow;
    double amp_re, amp_im;
    double dt;
};

__device__ inline void dipole_dev(const PulseDev& s, double t, double* re, double* im) {
    if (s.kind == 0) {
        const double tt = t - s.peak;
        if ((double)(float)fabs(tt) > s.cutoff) {
            *re = *im = 0.0;
            return;
        }
        const double g = exp(-tt * tt / (2 * s.width * s.width));
        double sn, cs;
        sincos(-s.omega * tt - s.phi, &sn, &cs);
        // polar(1, th) / (-i omega) = (cos th + i sin th) (i / omega)
        *re = -g * sn / s.omega;
        *im = g * cs / s.omega;
    } else {
        const float rt = (float)t;
        if (rt < s.t_start || rt > s.t_end) {
            *re = *im = 0.0;
            return;
        }
        const double w = 2 * M_PI * s.freq;
        double env = 1.0;
        if (s.cw_width != 0.0) {
            const double ts = (t - s.t_start) / s.cw_width - s.slow;
            const double te = (s.t_end - t) / s.cw_width - s.slow;
            env = (1.0 + tanh(ts)) * (1.0 + tanh(te)) * 0.25;
        }
        double sn, cs;
        sincos(-w * t, &sn, &cs);
        *re = -env * sn / w;
        *im = env * cs / w;
    }
}

// amp * J^{n+1/2}, J = (dipole(t_{n+1}) - dipole(t_n)) / dt, real or imaginary part
__device__ inline double source_current(const PulseDev& s, long step) {
    double r0, i0, r1, i1;
    dipole_dev(s, step * s.dt, &r0, &i0);
    dipole_dev(s, (step + 1) * s.dt, &r1, &i1);
    const double jr = (r1 - r0) / s.dt, ji = (i1 - i0) / s.dt;
    return s.part ? s.amp_re * ji + s.amp_im * jr : s.amp_re * jr - s.amp_im * ji;
}

// amp * K^n of a magnetic current, K^n = (dipole(t_n + dt/2) - dipole(t_n - dt/2)) / dt: centred on
// t_n, the midpoint of the H update from n - 1/2 to n + 1/2, as J^{n+1/2} is on that of the E update
__device__ inline double source_current_h(const PulseDev& s, long step) {
    double r0, i0, r1, i1;
    dipole_dev(s, (step - 0.5) * s.dt, &r0, &i0);
    dipole_dev(s, (step + 0.5) * s.dt, &r1, &i1);
    const double kr = (r1 - r0) / s.dt, ki = (i1 - i0) / s.dt;
    return s.part ? s.amp_re * ki + s.amp_im * kr : s.amp_re * kr - s.amp_im * ki;
}

// ---- source injection: E -= coef * J on the source's cells (coef = dt * weight * inv_eps)
// One block covers entries of one source; the block's source and entry range come from
// per-block tables, so all sources of all engines in a launch set go in one launch.
template <typename T>
struct SrcDev {
    PulseDev p;
    T* F[2];             // field set 0 / 1 (entries hold layout indices of the source component)
    const long* d_step;  // the owning engine's step counter
};

// the tables of the source launch: per workgroup its source and range of entries; per entry
// the layout index and FP64 coefficient
template <typename T>
struct SrcTables {
    const SrcDev<T>* src;
    const int* blk_src;
    const long* blk_first;
    const long* blk_end;
    const long* idx;
    const double* coef;
};

// the cells of source workgroup blk
template <typename T>
__device__ __forceinline__ void source_block(const SrcTables<T>& t, const int blk) {
    __shared__ double J;
    const int sid = t.blk_src[blk];
    const SrcDev<T>& S = t.src[sid];
    const long step = *S.d_step;
    if (threadIdx.x == 0) J = source_current(S.p, step);
    __syncthreads();
    const long e = t.blk_first[blk] + threadIdx.x;
    if (e >= t.blk_end[blk] || J == 0.0) return;
    T* F = S.F[(step + 1) & 1];  // the step just completed wrote the other buffer
    const long q = t.idx[e];
    F[q] = (T)((double)F[q] - t.coef[e] * J);
}

// Magnetic currents (dB/dt = -curl E - K, mu = 1): H^{n-1/2} -= dt * weight * K^n in the set the
// coming step reads, before the fused update; the update then forms
// H^{n+1/2} = (H^{n-1/2} - dt K^n) - dt curl E^n, so every workgroup that recomputes a halo row of H
// sees the injected value.  Launched after the monitor gather of step n (samples see H^{n-1/2}).
template <typename T>
__global__ void __launch_bounds__(256) k_source_h(const SrcTables<T> t) {
    __shared__ double K;
    const int blk = blockIdx.x;
    const int sid = t.blk_src[blk];
    const SrcDev<T>& S = t.src[sid];
    const long step = *S.d_step;
    if (threadIdx.x == 0) K = source_current_h(S.p, step);
    __syncthreads();
    const long e = t.blk_first[blk] + threadIdx.x;
    if (e >= t.blk_end[blk] || K == 0.0) return;
    T* F = S.F[step & 1];  // the set step n reads
    const long q = t.idx[e];
    F[q] = (T)((double)F[q] - t.coef[e] * K);
}

// A step with sources and no DFT region ends here: the source kernel advances the counters.
template <typename T>
__global__ void __launch_bounds__(256) k_source(const SrcTables<T> t, const Tick tk) {
    source_block(t, blockIdx.x);
    tick_tail(tk);
}

// ---- point monitors: trilinear gather, FP64 result --------------------------------------
template <typename T>
struct MonDev {
    const T* F[2];       // field set 0 / 1 (idx holds layout indices of the sampled component)
    const long* d_step;  // owning engine's counter
    double* out;         // sample buffer of the owning engine: [sample][npt]
    int npt;             // points of the owning engine
    int p;               // index of this point within its engine
    int span;            // save span
    long idx[8];
    double w[8];
};

template <typename T>
__global__ void __launch_bounds__(256) k_monitor(const MonDev<T>* __restrict__ mon, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MonDev<T>& m = mon[i];
    const long step = *m.d_step;
    if (step % m.span) return;
    const T* f = m.F[step & 1];
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (m.idx[q] >= 0) v += m.w[q] * (double)f[m.idx[q]];
    m.out[(step / m.span) * m.npt + m.p] = v;
}

// ---- DFT: Neumaier-compensated FP64 accumulation of F exp(i 2 pi f t) dt ----------------
__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

template <typename T>
struct DftDev {
    const T* F[2];       // field set 0 / 1 base, [x][6][y][z]
    const long* d_step;  // owning engine's counter
    long N, nyz, nz;
    int lo[3], ext[3];
    long ncell;
    long first;          // first global index of this region's (component, cell) pairs
    int comps[6];
    int ncomp;
    int nf;
    const double* freq;  // nf frequencies
    double dt;
    double2* acc;        // [slot][f][cell]
    double2* cmp;        // compensation terms, same layout
};

// accumulation of one (component, cell, frequency) of region ri
template <typename T>
__device__ __forceinline__ void dft_cell(const DftDev<T>* __restrict__ regs, int ri, int f, long g);

// One thread per (component, cell, frequency): blockIdx.y is the frequency index, and the
// threads of a region run over its components' cells one component after another.  The phase
// factors depend only on the step and the frequency; each thread evaluates them itself (two
// FP64 sincos), which measured faster than one thread per block sharing them through LDS
// behind a barrier.
template <typename T>
__device__ __forceinline__ void dft_block(const DftDev<T>* __restrict__ regs, const int nreg, const long total,
                                          const int bx, const int f) {
    const long g = (long)bx * blockDim.x + threadIdx.x;
    auto region_of = [&](long x) {
        int lo = 0, hi = nreg - 1;
        while (lo < hi) {  // region with first <= x
            const int mid = (lo + hi + 1) / 2;
            if (regs[mid].first <= x) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    if (g < total) dft_cell(regs, region_of(g), f, g);
}

template <typename T>
__global__ void __launch_bounds__(256) k_dft(const DftDev<T>* __restrict__ regs, int nreg, long total, const Tick tk) {
    dft_block(regs, nreg, total, blockIdx.x, blockIdx.y);
    tick_tail(tk);
}

// Source injection and DFT in one launch, for steps whose DFT regions hold no source cell (so
// neither part reads what the other writes): workgroups [0, nsrc) inject, the rest accumulate
// (dftx workgroups per frequency); the last workgroup advances the step counters.
template <typename T>
__global__ void __launch_bounds__(256) k_source_dft(const SrcTables<T> t, int nsrc, const DftDev<T>* __restrict__ regs,
                                                    int nreg, long total, int dftx, const Tick tk) {
    const int b = blockIdx.x;
    if (b < nsrc) {
        source_block(t, b);
    } else {
        const int d = b - nsrc;
        dft_block(regs, nreg, total, d % dftx, d / dftx);
    }
    tick_tail(tk);
}

template <typename T>
__device__ __forceinline__ void dft_cell(const DftDev<T>* __restrict__ regs, const int ri, const int f, const long g) {
    const DftDev<T>& a = regs[ri];
    if (f >= a.nf) return;
    const long local = g - a.first;
    const int s = (int)(local / a.ncell);  // component slot
    const long cell = local - (long)s * a.ncell;
    const long step = *a.d_step;
    const T* F = a.F[(step + 1) & 1];
    const int kk = (int)(cell % a.ext[2]);
    const long rest = cell / a.ext[2];
    const int jj = (int)(rest % a.ext[1]);
    const int ii = (int)(rest / a.ext[1]);
    const long q = (long)(a.lo[0] + ii) * a.nyz + (long)(a.lo[1] + jj) * a.nz + (a.lo[2] + kk);
    // phases of this thread's frequency at the E and H time levels, t = (n+1) dt and (n+1/2) dt
    // (each thread forms its own: no serial set-up ahead of the loads)
    double cE, sE, cH, sH;
    const double w = 2 * M_PI * a.freq[f];
    sincos(w * ((step + 1) * a.dt), &sE, &cE);
    sincos(w * ((step + 0.5) * a.dt), &sH, &cH);
    const int c = a.comps[s];
    const double v = (double)F[field_index(q, c, a.nyz)] * a.dt;
    const long o = ((long)s * a.nf + f) * a.ncell + cell;
    double2 sm = a.acc[o], cp = a.cmp[o];
    neumaier(sm.x, cp.x, v * (c < 3 ? cE : cH));
    neumaier(sm.y, cp.y, v * (c < 3 ? sE : sH));
    a.acc[o] = sm;
    a.cmp[o] = cp;
}

// ---- flux through a DFT plane ------------------------------------------------------------
// out[f] = dA * Re sum_cells (E_a conj(H_b) - E_b conj(H_a)) for each frequency f, with each
// spectrum value acc + cmp of the region's [slot][f][cell] accumulators (sa, sb, ha, hb: the slots
// of E_a, E_b, H_a, H_b).  Two launches, both in a fixed order, so the result does not vary from
// run to run:
//   k_dft_flux_part, grid (nb, nf): workgroup (b, f) sums cells b*256 + t + m*nb*256 in FP64 per
//     thread, the wavefronts reduce by shuffles and the first thread adds the wavefront sums in
//     order into part[f * nb + b];
//   k_dft_flux_sum, one thread per frequency: out[f] = dA * (part[f*nb] + ... + part[f*nb + nb-1]).
__global__ void __launch_bounds__(256) k_dft_flux_part(const double2* __restrict__ acc, const double2* __restrict__ cmp,
                                                       long ncell, int nf, int sa, int sb, int ha, int hb, int nb,
                                                       double* __restrict__ part) {
    const int b = blockIdx.x, f = blockIdx.y;
    auto value = [&](int slot, long c) {
        const long o = ((long)slot * nf + f) * ncell + c;
        const double2 x = acc[o], y = cmp[o];
        return make_double2(x.x + y.x, x.y + y.y);
    };
    double sum = 0;
    const long stride = (long)nb * blockDim.x;
    for (long c = (long)b * blockDim.x + threadIdx.x; c < ncell; c += stride) {
        const double2 ea = value(sa, c), eb = value(sb, c), Ha = value(ha, c), Hb = value(hb, c);
        sum += (ea.x * Hb.x + ea.y * Hb.y) - (eb.x * Ha.x + eb.y * Ha.y);
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    __shared__ double wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += wsum[w];
        part[(long)f * nb + b] = t;
    }
}

__global__ void __launch_bounds__(64) k_dft_flux_sum(const double* __restrict__ part, int nf, int nb, double dA,
                                                     double* __restrict__ out) {
    const int f = threadIdx.x;
    if (f >= nf) return;
    double t = 0;
    for (int b = 0; b < nb; ++b) t += part[(long)f * nb + b];
    out[f] = t * dA;
}

// ===================================================== adjoint (reverse-mode) kernels ======
// Exact transpose of one k_step update (Yee curls, CPML recursions, (P, Q) pole recursions),
// applied backwards in time to the adjoint state (Ebar, Hbar, Qbar, Pbar, psiEbar, psiHbar).
// With b the CPML decay of a derivative term and a = b - 1, the forward term
//   d~ = b (d + psi),  psi' = b psi + (b - 1) d
// transposes to  dbar = b dbar~ + (b - 1) psibar',  psibar = b (dbar~ + psibar').
// Non-PML terms have b = 1.  The adjoint of a backward difference e(r) = H(r) - H(r-1) adds
// ebar(q) - ebar(q+1) to Hbar(q); the adjoint of a forward difference d(r) = E(r+1) - E(r)
// adds dbar(q-1) - dbar(q) to Ebar(q); positions outside a metallic edge contribute zero.
// One fused launch per reverse step (adj_tile) does both halves:
//   phase 1, Hbar_total(q) = Hbar'(q) + sum over E-terms [ebar(q) - ebar(q + e_d)], and the new
//            H auxiliary bars psiHbar = bH (-s cdt Hbar_total + psiHbar');
//   phase 2, Ebar(q) = Ebar'(q) + sum over H-terms [dbar(q - e_d) - dbar(q)] + the pole
//            recursions, and the new E auxiliary bars psiEbar = bE (s cdt u + psiEbar');
// with u = inv_eps Ebar', ebar = s cdt u (PML: bE s cdt u + (bE - 1) psiEbar'), dbar = -s cdt
// Hbar_total (PML: bH (-s cdt Hbar_total) + (bH - 1) psiHbar').  The terms, as (E comp, axis,
// sign, H comp, aux slot on that axis):
//   E-terms: (Ex, y, +, Hz, 0) (Ex, z, -, Hy, 0) (Ey, z, +, Hx, 1) (Ey, x, -, Hz, 0)
//            (Ez, x, +, Hy, 1) (Ez, y, -, Hx, 1)
//   H-terms: (Hx, y, +, Ez, 0) (Hx, z, -, Ey, 0) (Hy, z, +, Ex, 1) (Hy, x, -, Ez, 0)
//            (Hz, x, +, Ey, 1) (Hz, y, -, Ex, 1)
// This has the dependency pattern of the forward step (phase 1 uses forward differences,
// phase 2 backward ones), so the same tiling applies: a workgroup spans a z line, each thread
// one y row plus a halo row below it where phase 1 is recomputed, z neighbours through LDS,
// and the tile marches along x carrying u of the next planes.  Adjoint fields and E auxiliary
// bars are double-buffered (neighbouring workgroups read the old values of the halo row);
// Qbar and Pbar are only touched by their owner and are updated in place.

template <typename T>
struct AdjArgs {
    StepArgs<T> f;     // forward description (geometry, materials, PML tables, tiles)
    T* A[2];           // adjoint field sets [x][Ubar x,y,z, Hbar x,y,z][y][z] with Ubar = Ebar / eps:
                       // read A[rd], write A[1 - rd]
    int rd;
    T* Aq;             // Qbar [x][slot * 3 + comp][y][z]
    T* Ap;             // Pbar, the same layout
    T* Apsi[3];        // per axis [8][S]: E aux bar pair (0,1), H aux bar pairs (2,3) and (4,5),
                       // E aux bar pair (6,7)
    int er, ew;        // E aux bar pair read (0 or 6) / written (6 or 0)
    int hr, hw;        // H aux bar pair read (2 or 4) / written (4 or 2)
};

// One tile of the fused reverse step: a z line of one y row (j0), marching along x.  Addressing
// and predication as in step_tile: per-plane descriptors, lane offsets fixed for the march, the
// out-of-range offset for ghost rows and for lanes past the z line, outside a z PML layer, a
// pole slot or the design region; whole-wavefront conditions are uniform branches.
template <typename T, int NP, bool NTS>
__device__ __forceinline__ void adj_tile(const AdjArgs<T>& g, const int t) {
    constexpr uint32_t TS = sizeof(T);
    const StepArgs<T>& a = g.f;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int B = blockDim.x;
    T* se = reinterpret_cast<T*>(smem_raw);  // [2 parity][4][B]: z E-terms of rows -1 and 0
    T* sd = se + 8 * B;                      // [2 parity][2][B]: z H-terms of row 0
    MatLut<T>* slut = reinterpret_cast<MatLut<T>*>(sd + 4 * B);
    {
        GLOBAL const uint32_t* src = (GLOBAL const uint32_t*)a.lut;
        uint32_t* dst = reinterpret_cast<uint32_t*>(slut);
        const int words = (int)(sizeof(MatLut<T>) / 4) * a.nmat;
        for (int q = threadIdx.x; q < words; q += B) dst[q] = src[q];
    }
    const char* AR = (const char*)g.A[g.rd];
    const char* AW = (const char*)g.A[1 - g.rd];
    const char* matw = (const char*)a.matw;
    GLOBAL const T* coef = (GLOBAL const T*)a.coef;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int j0 = t % a.nty, xci = t / a.nty;
    const int x0 = xci * a.xc, x1 = min(x0 + a.xc, nx);
    const int k = threadIdx.x;
    const bool kin = k < nz;
    const int kp1 = wrap_or_ghost(k + 1, nz, a.pz), km1 = wrap_or_ghost(k - 1, nz, a.pz);
    const T cdt = a.cdt;
    const bool uniform = a.uniform != 0;
    const uint32_t er = g.er, ew = g.ew, hr = g.hr, hw = g.hw;
    const uint32_t PB = (uint32_t)ny * (uint32_t)nz * TS;  // bytes of one component plane
    const long FPB = 6L * PB, QPB = (long)a.nqp * PB, PW = (long)ny * nz * 4;
    const int Ly = a.plo[1] + a.phi[1], Lz = a.plo[2] + a.phi[2];
    const uint32_t SB0 = (uint32_t)(a.S[0] * TS), SB1 = (uint32_t)(a.S[1] * TS), SB2 = (uint32_t)(a.S[2] * TS);
    const Rsrc rpx = rsrc(g.Apsi[0]), rpy = rsrc(g.Apsi[1]), rpz = rsrc(g.Apsi[2]);
    // rows: 0 = j0 - 1 (halo), 1 = j0 (owned), 2 = j0 + 1 (its +y neighbour); uniform validity
    int gj[3], sy[3];
    bool rv[3], ypml[3];
    uint32_t oR[3], oW[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        gj[r] = wrap_or_ghost(j0 + r - 1, ny, a.py);
        rv[r] = gj[r] >= 0;
        ypml[r] = rv[r] && in_pml(a, 1, gj[r], ny);
        sy[r] = ypml[r] ? (gj[r] < a.plo[1] ? gj[r] : gj[r] - (ny - a.phi[1]) + a.plo[1]) : 0;
        const bool ok = rv[r] && kin;
        oR[r] = ok ? (uint32_t)(gj[r] * nz + k) * TS : kOOB;
        oW[r] = ok ? (uint32_t)(gj[r] * nz + k) * 4u : kOOB;
    }
    const bool zpml = kin && in_pml(a, 2, k, nz);
    const bool zany = wave_any(zpml);
    const uint32_t oK = kin ? (uint32_t)k * TS : kOOB;
    const uint32_t oZ = zpml ? (uint32_t)(k < a.plo[2] ? k : k - (nz - a.phi[2]) + a.plo[2]) * TS : kOOB;
    GLOBAL const T* cy = coef + a.coff[1];
    GLOBAL const T* cz = coef + a.coff[2];
    GLOBAL const T* cx = coef + a.coff[0];
    const T bEz = zpml ? cz[k] : T(1), bHz = zpml ? cz[2 * nz + k] : T(1);
    const Rsrc rd = rsrc(a.dinv);
    const int dey = a.dhi[1] - a.dlo[1], dez = a.dhi[2] - a.dlo[2];

    const int istart = (x0 > 0 || a.px) ? x0 - 1 : x0;
    // u = Ubar' (the state itself) on planes i and i + 1 (rows 0..2; plane i + 1 is loaded with
    // the other loads of iteration i) and the owned row's material word of both planes
    T uc[3][3], un[3][3];
    uint32_t wc1 = 0u, wn1 = 0u;
    auto load_u = [&](int p, T (&u)[3][3], uint32_t& w1) {
        const int pp = wrap_or_ghost(p, nx, a.px);
        const long pc = pp >= 0 ? pp : 0;
        const Rsrc rf = rsrc(AR + pc * FPB);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const uint32_t o = pp >= 0 ? oR[r] : kOOB;
#pragma unroll
            for (int c = 0; c < 3; ++c) u[r][c] = bld<T>(rf, o, c * PB);
        }
        w1 = uniform ? 0u : bld<uint32_t>(rsrc(matw + pc * PW), pp >= 0 ? oW[1] : kOOB, 0);
    };
    lds_barrier();  // LUT in LDS
    load_u(istart, uc, wc1);
    T dbx_prev[2] = {T(0), T(0)};  // x H-terms of row 0 on plane i - 1: (Hz -> Ey), (Hy -> Ez)

    for (int i = istart; i < x1; ++i) {
        const bool own_plane = i >= x0;
        const int pi = wrap_or_ghost(i, nx, a.px);
        const int pn = wrap_or_ghost(i + 1, nx, a.px);
        load_u(i + 1, un, wn1);
        const bool xpml = in_pml(a, 0, pi, nx);
        const bool xpmln = pn >= 0 && in_pml(a, 0, pn, nx);
        const uint32_t sxo = xpml ? (uint32_t)(pi < a.plo[0] ? pi : pi - (nx - a.phi[0]) + a.plo[0]) * PB : 0u;
        const uint32_t sxn = xpmln ? (uint32_t)(pn < a.plo[0] ? pn : pn - (nx - a.phi[0]) + a.plo[0]) * PB : 0u;
        const T bEx = xpml ? cx[pi] : T(1), bHx = xpml ? cx[2 * nx + pi] : T(1);
        const T bExn = xpmln ? cx[pn] : T(1);
        const Rsrc rA = rsrc(AR + (long)pi * FPB);
        const Rsrc rB = rsrc(AW + (long)pi * FPB);

        // ---- the global loads of plane i, issued together (u of the planes ahead is already in
        // flight): Hbar' of rows -1 and 0, the z auxiliary bars, the owned cell's design 1/eps
        // and pole state; the y and x auxiliary bars (layers of a few tiles) are loaded in place
        T hb[2][3];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) hb[r][c] = bld<T>(rA, oR[r], (3 + c) * PB);
        T pzE[2][2] = {{T(0), T(0)}, {T(0), T(0)}}, phz[2] = {T(0), T(0)};
        if (zany) {  // lanes outside the layers: b = 1 and a zero auxiliary
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                if (!rv[r]) continue;
                const uint32_t so = (uint32_t)((pi * ny + gj[r]) * Lz) * TS;
                pzE[r][0] = bld<T>(rpz, oZ, er * SB2 + so);
                pzE[r][1] = bld<T>(rpz, oZ, (er + 1) * SB2 + so);
                if (r == 1) {
                    phz[0] = bld<T>(rpz, oZ, hr * SB2 + so);
                    phz[1] = bld<T>(rpz, oZ, (hr + 1) * SB2 + so);
                }
            }
        }
        const bool wr = own_plane && rv[1];
        // material of the owned cell: pole count per component and the LUT entry (design voxels:
        // 1/eps from the design array, no poles)
        int mm[3] = {0, 0, 0}, npl[3] = {0, 0, 0};
        bool des[3] = {false, false, false};
        T dv[3];
        bool desany = false;
        uint32_t sact = 0u;  // pole slots some lane of the wavefront uses (uniform)
        T qv[NP > 0 ? NP : 1][3], pv[NP > 0 ? NP : 1][3];
        if (wr && !uniform) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int m = (wc1 >> (8 * c)) & 0xff;
                des[c] = m == kDesignMat;
                mm[c] = des[c] ? 0 : m;
                npl[c] = des[c] ? 0 : slut[mm[c]].np;
            }
            desany = wave_any(des[0] || des[1] || des[2]);
            if (desany) {
                const uint32_t dq =
                    (uint32_t)(((pi - a.dlo[0]) * dey + (gj[1] - a.dlo[1])) * dez + (k - a.dlo[2])) * TS;
#pragma unroll
                for (int c = 0; c < 3; ++c) dv[c] = bld<T>(rd, des[c] ? dq : kOOB, 0);
            }
            if (NP > 0) {
                const uint32_t o = oR[1];
                const Rsrc rQ = rsrc((const char*)g.Aq + (long)pi * QPB);
                const Rsrc rP = rsrc((const char*)g.Ap + (long)pi * QPB);
#pragma unroll
                for (int s = 0; s < NP; ++s) {
                    if (!wave_any(s < npl[0] || s < npl[1] || s < npl[2])) continue;
                    sact |= 1u << s;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const bool vq = s < npl[c];
                        const bool vp = vq && ((slut[mm[c]].lorentz_mask >> s) & 1);
                        const uint32_t so = (s * 3 + c) * PB;
                        pv[s][c] = bld<T>(rP, vp ? o : kOOB, so);
                        qv[s][c] = bld<T>(rQ, vq ? o : kOOB, so);
                    }
                }
            }
        }

        // ---- phase 1: E-terms and Hbar_total on rows -1 and 0 of plane i ----------------
        // E-terms on rows r = 0, 1 (labels: y (Ex->Hz), z (Ex->Hy), z (Ey->Hx), x (Ey->Hz),
        // x (Ez->Hy), y (Ez->Hx)) and the y-terms of row 2
        T eb[3][6];
        T pey[2] = {T(0), T(0)}, pex[2] = {T(0), T(0)};  // owned row's y and x E aux bars
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const bool ok = rv[r];
            const int j = gj[r];
            const T ux = uc[r][0], uy = uc[r][1], uz = uc[r][2];
            T t_ex_y = cdt * ux, t_ex_z = -cdt * ux, t_ey_z = cdt * uy, t_ey_x = -cdt * uy, t_ez_x = cdt * uz,
              t_ez_y = -cdt * uz;
            if (ypml[r]) {
                const T b = cy[j];
                const uint32_t so = (uint32_t)((pi * Ly + sy[r]) * nz) * TS;
                const T p0 = bld<T>(rpy, oK, er * SB1 + so), p1 = bld<T>(rpy, oK, (er + 1) * SB1 + so);
                t_ex_y = b * t_ex_y + (b - T(1)) * p0;
                t_ez_y = b * t_ez_y + (b - T(1)) * p1;
                if (r == 1) {
                    pey[0] = p0;
                    pey[1] = p1;
                }
            }
            if (r < 2) {
                if (ok && zany) {
                    t_ex_z = bEz * t_ex_z + (bEz - T(1)) * pzE[r][0];
                    t_ey_z = bEz * t_ey_z + (bEz - T(1)) * pzE[r][1];
                }
                if (ok && xpml) {
                    const T p0 = bld<T>(rpx, oR[r], er * SB0 + sxo), p1 = bld<T>(rpx, oR[r], (er + 1) * SB0 + sxo);
                    t_ey_x = bEx * t_ey_x + (bEx - T(1)) * p0;
                    t_ez_x = bEx * t_ez_x + (bEx - T(1)) * p1;
                    if (r == 1) {
                        pex[0] = p0;
                        pex[1] = p1;
                    }
                }
            }
            eb[r][0] = t_ex_y;
            eb[r][1] = t_ex_z;
            eb[r][2] = t_ey_z;
            eb[r][3] = t_ey_x;
            eb[r][4] = t_ez_x;
            eb[r][5] = t_ez_y;
        }
        // x-terms of plane i + 1, rows 0 and 1
        T ebn[2][2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            T a1 = -cdt * un[r][1], a2 = cdt * un[r][2];  // (Ey -> Hz), (Ez -> Hy)
            if (rv[r] && xpmln) {
                a1 = bExn * a1 + (bExn - T(1)) * bld<T>(rpx, oR[r], er * SB0 + sxn);
                a2 = bExn * a2 + (bExn - T(1)) * bld<T>(rpx, oR[r], (er + 1) * SB0 + sxn);
            }
            ebn[r][0] = (rv[r] && pn >= 0) ? a1 : T(0);
            ebn[r][1] = (rv[r] && pn >= 0) ? a2 : T(0);
        }
        // z-terms through LDS
        T* bufe = se + (i & 1) * 4 * B;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            bufe[(2 * r + 0) * B + k] = eb[r][1];  // (Ex -> Hy)
            bufe[(2 * r + 1) * B + k] = eb[r][2];  // (Ey -> Hx)
        }
        lds_barrier();
        T tot[2][3];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const bool ok = rv[r];
            const T exz_k = kp1 >= 0 ? bufe[(2 * r + 0) * B + kp1] : T(0);
            const T eyz_k = kp1 >= 0 ? bufe[(2 * r + 1) * B + kp1] : T(0);
            // Hx: (Ey, z) and (Ez, y); Hy: (Ex, z) and (Ez, x); Hz: (Ex, y) and (Ey, x)
            tot[r][0] = ok ? hb[r][0] + (eb[r][2] - eyz_k) + (eb[r][5] - eb[r + 1][5]) : T(0);
            tot[r][1] = ok ? hb[r][1] + (eb[r][1] - exz_k) + (eb[r][4] - ebn[r][1]) : T(0);
            tot[r][2] = ok ? hb[r][2] + (eb[r][0] - eb[r + 1][0]) + (eb[r][3] - ebn[r][0]) : T(0);
        }
        // H auxiliary bars of the owned row (read by the update below and by phase 2)
        T phy[2] = {T(0), T(0)}, phx[2] = {T(0), T(0)};
        if (ypml[1]) {
            const uint32_t so = (uint32_t)((pi * Ly + sy[1]) * nz) * TS;
            phy[0] = bld<T>(rpy, oK, hr * SB1 + so);
            phy[1] = bld<T>(rpy, oK, (hr + 1) * SB1 + so);
        }
        if (rv[1] && xpml) {
            phx[0] = bld<T>(rpx, oR[1], hr * SB0 + sxo);
            phx[1] = bld<T>(rpx, oR[1], (hr + 1) * SB0 + sxo);
        }
        // own row: Hbar at the old time and the new H auxiliary bars
        if (wr) {
            const int j = gj[1];
            bst<NTS>(rB, oR[1], 3 * PB, tot[1][0]);
            bst<NTS>(rB, oR[1], 4 * PB, tot[1][1]);
            bst<NTS>(rB, oR[1], 5 * PB, tot[1][2]);
            if (ypml[1]) {  // (Hx, y, +, slot 0), (Hz, y, -, slot 1)
                const T b = cy[2 * ny + j];
                const uint32_t so = (uint32_t)((pi * Ly + sy[1]) * nz) * TS;
                bst<NTS>(rpy, oK, hw * SB1 + so, b * (-cdt * tot[1][0] + phy[0]));
                bst<NTS>(rpy, oK, (hw + 1) * SB1 + so, b * (cdt * tot[1][2] + phy[1]));
            }
            if (zany) {  // (Hx, z, -, slot 0), (Hy, z, +, slot 1)
                const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
                bst<NTS>(rpz, oZ, hw * SB2 + so, bHz * (cdt * tot[1][0] + phz[0]));
                bst<NTS>(rpz, oZ, (hw + 1) * SB2 + so, bHz * (-cdt * tot[1][1] + phz[1]));
            }
            if (xpml) {  // (Hy, x, -, slot 0), (Hz, x, +, slot 1)
                bst<NTS>(rpx, oR[1], hw * SB0 + sxo, bHx * (cdt * tot[1][1] + phx[0]));
                bst<NTS>(rpx, oR[1], (hw + 1) * SB0 + sxo, bHx * (-cdt * tot[1][2] + phx[1]));
            }
        }

        // ---- phase 2: H-terms and Ebar on row 0 of plane i --------------------------------
        // H-terms of row 0: (Hx, y) (Hx, z) (Hy, z) (Hy, x) (Hz, x) (Hz, y); y-terms of row -1
        T db[6], dbh[2];
        {
            const int j = gj[1];
            T d_hx_y = -cdt * tot[1][0], d_hx_z = cdt * tot[1][0], d_hy_z = -cdt * tot[1][1], d_hy_x = cdt * tot[1][1],
              d_hz_x = -cdt * tot[1][2], d_hz_y = cdt * tot[1][2];
            if (ypml[1]) {
                const T b = cy[2 * ny + j];
                d_hx_y = b * d_hx_y + (b - T(1)) * phy[0];
                d_hz_y = b * d_hz_y + (b - T(1)) * phy[1];
            }
            if (rv[1] && zany) {
                d_hx_z = bHz * d_hx_z + (bHz - T(1)) * phz[0];
                d_hy_z = bHz * d_hy_z + (bHz - T(1)) * phz[1];
            }
            if (rv[1] && xpml) {
                d_hy_x = bHx * d_hy_x + (bHx - T(1)) * phx[0];
                d_hz_x = bHx * d_hz_x + (bHx - T(1)) * phx[1];
            }
            db[0] = d_hx_y;
            db[1] = d_hx_z;
            db[2] = d_hy_z;
            db[3] = d_hy_x;
            db[4] = d_hz_x;
            db[5] = d_hz_y;
        }
        {
            const int j = gj[0];
            T d_hx_y = -cdt * tot[0][0], d_hz_y = cdt * tot[0][2];
            if (ypml[0]) {
                const T b = cy[2 * ny + j];
                const uint32_t so = (uint32_t)((pi * Ly + sy[0]) * nz) * TS;
                d_hx_y = b * d_hx_y + (b - T(1)) * bld<T>(rpy, oK, hr * SB1 + so);
                d_hz_y = b * d_hz_y + (b - T(1)) * bld<T>(rpy, oK, (hr + 1) * SB1 + so);
            }
            dbh[0] = rv[0] ? d_hx_y : T(0);
            dbh[1] = rv[0] ? d_hz_y : T(0);
        }
        if (own_plane) {
            T* bufd = sd + (i & 1) * 2 * B;
            bufd[0 * B + k] = db[2];  // (Hy -> Ex)
            bufd[1 * B + k] = db[1];  // (Hx -> Ey)
            lds_barrier();
            if (wr) {
                const int j = gj[1];
                const uint32_t o = oR[1];
                const T hyz_k = km1 >= 0 ? bufd[0 * B + km1] : T(0);
                const T hxz_k = km1 >= 0 ? bufd[1 * B + km1] : T(0);
                // Ex: (Hy, z) and (Hz, y); Ey: (Hx, z) and (Hz, x); Ez: (Hx, y) and (Hy, x)
                const T u0[3] = {uc[1][0], uc[1][1], uc[1][2]};
                // 1/eps and eps per component of the owned cell
                T inv[3], epsc[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    inv[c] = slut[mm[c]].inv_eps;
                    epsc[c] = slut[mm[c]].eps_inf;
                }
                if (desany) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        inv[c] = des[c] ? dv[c] : inv[c];
                        epsc[c] = des[c] ? T(1) / dv[c] : epsc[c];
                    }
                }
                // Ebar' = eps Ubar' plus the differences
                T out[3];
                out[0] = epsc[0] * u0[0] + (hyz_k - db[2]) + (dbh[1] - db[5]);
                out[1] = epsc[1] * u0[1] + (hxz_k - db[1]) + (dbx_prev[0] - db[4]);
                out[2] = epsc[2] * u0[2] + (dbh[0] - db[0]) + (dbx_prev[1] - db[3]);
                if (NP > 0 && !uniform) {
                    const Rsrc rQ = rsrc((const char*)g.Aq + (long)pi * QPB);
                    const Rsrc rP = rsrc((const char*)g.Ap + (long)pi * QPB);
#pragma unroll
                    for (int s = 0; s < NP; ++s) {
                        if (!((sact >> s) & 1u)) continue;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const MatLut<T>& L = slut[mm[c]];
                            const bool vq = s < npl[c];
                            const bool vp = vq && ((L.lorentz_mask >> s) & 1);
                            const uint32_t so = (s * 3 + c) * PB;
                            const T pb = pv[s][c];
                            const T qb = vq ? qv[s][c] + pb - u0[c] : T(0);
                            bst<NTS>(rQ, vq ? o : kOOB, so, L.alpha[s] * qb);
                            bst<NTS>(rP, vp ? o : kOOB, so, pb + L.beta[s] * qb);
                            out[c] += L.c3[s] * qb;
                        }
                    }
                }
                // new E auxiliary bars: E-terms of row 0 with their signs
                if (ypml[1]) {  // (Ex, y, +, slot 0), (Ez, y, -, slot 1)
                    const T b = cy[j];
                    const uint32_t so = (uint32_t)((pi * Ly + sy[1]) * nz) * TS;
                    bst<NTS>(rpy, oK, ew * SB1 + so, b * (cdt * u0[0] + pey[0]));
                    bst<NTS>(rpy, oK, (ew + 1) * SB1 + so, b * (-cdt * u0[2] + pey[1]));
                }
                if (zany) {  // (Ex, z, -, slot 0), (Ey, z, +, slot 1)
                    const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
                    bst<NTS>(rpz, oZ, ew * SB2 + so, bEz * (-cdt * u0[0] + pzE[1][0]));
                    bst<NTS>(rpz, oZ, (ew + 1) * SB2 + so, bEz * (cdt * u0[1] + pzE[1][1]));
                }
                if (xpml) {  // (Ey, x, -, slot 0), (Ez, x, +, slot 1)
                    bst<NTS>(rpx, o, ew * SB0 + sxo, bEx * (-cdt * u0[1] + pex[0]));
                    bst<NTS>(rpx, o, (ew + 1) * SB0 + sxo, bEx * (cdt * u0[2] + pex[1]));
                }
                // the state is Ubar = Ebar / eps
                bst<NTS>(rB, o, 0, inv[0] * out[0]);
                bst<NTS>(rB, o, PB, inv[1] * out[1]);
                bst<NTS>(rB, o, 2 * PB, inv[2] * out[2]);
            }
        } else {
            lds_barrier();  // keep the barrier count uniform across iterations
        }
        dbx_prev[0] = rv[1] ? db[4] : T(0);
        dbx_prev[1] = rv[1] ? db[3] : T(0);
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) uc[r][c] = un[r][c];
        wc1 = wn1;
    }
}

// One tile of the fused reverse step, row-of-threads form.  A workgroup holds a.rt rows of nz
// threads (one cell of a y row per thread): thread row 0 is the row below the tile, rows
// 1..rt-1 own rows j0..j0+rt-2, and the top thread row also forms the y-terms of the row above
// the tile.  The y and z neighbours' E-terms (for Hbar) and H-terms (for Ebar) pass through LDS,
// so each adjoint value is loaded once per tile and the y halo costs one row per rt-1 owned
// rows.  The arithmetic of every cell is that of adj_tile.
template <typename T, int NP>
__device__ __forceinline__ void adj_rows(const AdjArgs<T>& g, const int t) {
    constexpr uint32_t TS = sizeof(T);
    const StepArgs<T>& a = g.f;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int B = blockDim.x;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int RT = a.rt, R = RT - 1;
    T* sE = reinterpret_cast<T*>(smem_raw);  // [4][B]: E-terms z (Ex->Hy), z (Ey->Hx), y (Ex->Hz), y (Ez->Hx)
    T* sEt = sE + 4 * B;                     // [2][nz]: y-terms of the row above the tile
    T* sD = sEt + 2 * nz;                    // [4][B]: H-terms z (Hy->Ex), z (Hx->Ey), y (Hx->Ez), y (Hz->Ex)
    MatLut<T>* slut = reinterpret_cast<MatLut<T>*>(sD + 4 * B);
    {
        GLOBAL const uint32_t* src = (GLOBAL const uint32_t*)a.lut;
        uint32_t* dst = reinterpret_cast<uint32_t*>(slut);
        const int words = (int)(sizeof(MatLut<T>) / 4) * a.nmat;
        for (int q = threadIdx.x; q < words; q += B) dst[q] = src[q];
    }
    const char* AR = (const char*)g.A[g.rd];
    const char* AW = (const char*)g.A[1 - g.rd];
    const char* matw = (const char*)a.matw;
    GLOBAL const T* coef = (GLOBAL const T*)a.coef;
    const int tid = threadIdx.x;
    const bool act = tid < RT * nz;
    const int r = act ? tid / nz : 0;
    const int k = act ? tid - r * nz : 0;
    const int tyi = t % a.nty, xci = t / a.nty;
    const int x0 = xci * a.xc, x1 = min(x0 + a.xc, nx);
    const int jr = tyi * R + r - 1;  // unwrapped global row of this thread row
    const int j = act ? wrap_or_ghost(jr, ny, a.py) : -1;
    const bool rv = j >= 0;
    const bool owned = act && r >= 1 && jr < ny;  // implies rv
    const bool top = act && r == RT - 1;
    const int jt = top ? wrap_or_ghost(jr + 1, ny, a.py) : -1;
    const int kp1 = wrap_or_ghost(k + 1, nz, a.pz), km1 = wrap_or_ghost(k - 1, nz, a.pz);
    const T cdt = a.cdt;
    const bool uniform = a.uniform != 0;
    const uint32_t er = g.er, ew = g.ew, hr = g.hr, hw = g.hw;
    const uint32_t PB = (uint32_t)ny * (uint32_t)nz * TS;  // bytes of one component plane
    const long FPB = 6L * PB, QPB = (long)a.nqp * PB, PW = (long)ny * nz * 4;
    const int Ly = a.plo[1] + a.phi[1], Lz = a.plo[2] + a.phi[2];
    const uint32_t SB0 = (uint32_t)(a.S[0] * TS), SB1 = (uint32_t)(a.S[1] * TS), SB2 = (uint32_t)(a.S[2] * TS);
    const Rsrc rpx = rsrc(g.Apsi[0]), rpy = rsrc(g.Apsi[1]), rpz = rsrc(g.Apsi[2]);
    const bool ypml = rv && in_pml(a, 1, j, ny);
    const int sy = ypml ? (j < a.plo[1] ? j : j - (ny - a.phi[1]) + a.plo[1]) : 0;
    const bool ypmlt = jt >= 0 && in_pml(a, 1, jt, ny);
    const int syt = ypmlt ? (jt < a.plo[1] ? jt : jt - (ny - a.phi[1]) + a.plo[1]) : 0;
    const uint32_t oR = rv ? (uint32_t)(j * nz + k) * TS : kOOB;
    const uint32_t oW = owned ? (uint32_t)(j * nz + k) * 4u : kOOB;
    const uint32_t oT = jt >= 0 ? (uint32_t)(jt * nz + k) * TS : kOOB;
    const bool zpml = act && in_pml(a, 2, k, nz);
    const bool zany = wave_any(zpml);
    const uint32_t oK = act ? (uint32_t)k * TS : kOOB;
    const uint32_t oZ = zpml ? (uint32_t)(k < a.plo[2] ? k : k - (nz - a.phi[2]) + a.plo[2]) * TS : kOOB;
    GLOBAL const T* cy = coef + a.coff[1];
    GLOBAL const T* cz = coef + a.coff[2];
    GLOBAL const T* cx = coef + a.coff[0];
    const T bEz = zpml ? cz[k] : T(1), bHz = zpml ? cz[2 * nz + k] : T(1);
    const T bEy = ypml ? cy[j] : T(1), bHy = ypml ? cy[2 * ny + j] : T(1);
    const T bEyt = ypmlt ? cy[jt] : T(1);
    const Rsrc rd = rsrc(a.dinv);
    const int dey = a.dhi[1] - a.dlo[1], dez = a.dhi[2] - a.dlo[2];

    const int istart = (x0 > 0 || a.px) ? x0 - 1 : x0;
    // u = Ubar' of this thread's cell on planes i and i + 1, the top row's (Ubar'x, Ubar'z) of
    // the row above, and the owned cell's material word
    T uc[3], un[3], tc[2], tn[2];
    uint32_t wc = 0u, wn = 0u;
    auto load_u = [&](int p, T (&u)[3], T (&ut)[2], uint32_t& w) {
        const int pp = wrap_or_ghost(p, nx, a.px);
        const long pc = pp >= 0 ? pp : 0;
        const Rsrc rf = rsrc(AR + pc * FPB);
        const uint32_t o = pp >= 0 ? oR : kOOB;
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = bld<T>(rf, o, c * PB);
        ut[0] = ut[1] = T(0);
        if (top) {
            const uint32_t ot = pp >= 0 ? oT : kOOB;
            ut[0] = bld<T>(rf, ot, 0);
            ut[1] = bld<T>(rf, ot, 2 * PB);
        }
        w = uniform ? 0u : bld<uint32_t>(rsrc(matw + pc * PW), pp >= 0 ? oW : kOOB, 0);
    };
    lds_barrier();  // LUT in LDS
    load_u(istart, uc, tc, wc);
    T dbx_prev[2] = {T(0), T(0)};  // x H-terms of this row on plane i - 1: (Hz -> Ey), (Hy -> Ez)

    for (int i = istart; i < x1; ++i) {
        const bool own_plane = i >= x0;
        const int pi = wrap_or_ghost(i, nx, a.px);
        const int pn = wrap_or_ghost(i + 1, nx, a.px);
        load_u(i + 1, un, tn, wn);
        const bool xpml = in_pml(a, 0, pi, nx);
        const bool xpmln = pn >= 0 && in_pml(a, 0, pn, nx);
        const uint32_t sxo = xpml ? (uint32_t)(pi < a.plo[0] ? pi : pi - (nx - a.phi[0]) + a.plo[0]) * PB : 0u;
        const uint32_t sxn = xpmln ? (uint32_t)(pn < a.plo[0] ? pn : pn - (nx - a.phi[0]) + a.plo[0]) * PB : 0u;
        const T bEx = xpml ? cx[pi] : T(1), bHx = xpml ? cx[2 * nx + pi] : T(1);
        const T bExn = xpmln ? cx[pn] : T(1);
        const Rsrc rA = rsrc(AR + (long)pi * FPB);
        const Rsrc rB = rsrc(AW + (long)pi * FPB);

        // ---- the global loads of plane i: Hbar', the E and H auxiliary bars of this row, the
        // owned cell's design 1/eps and pole state
        T hb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) hb[c] = bld<T>(rA, oR, (3 + c) * PB);
        T pzE[2] = {T(0), T(0)}, phz[2] = {T(0), T(0)};
        if (zany && rv) {  // lanes outside the layers: b = 1 and a zero auxiliary
            const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
            pzE[0] = bld<T>(rpz, oZ, er * SB2 + so);
            pzE[1] = bld<T>(rpz, oZ, (er + 1) * SB2 + so);
            if (owned) {
                phz[0] = bld<T>(rpz, oZ, hr * SB2 + so);
                phz[1] = bld<T>(rpz, oZ, (hr + 1) * SB2 + so);
            }
        }
        T pey[2] = {T(0), T(0)}, phy[2] = {T(0), T(0)};
        if (ypml) {
            const uint32_t so = (uint32_t)((pi * Ly + sy) * nz) * TS;
            pey[0] = bld<T>(rpy, oK, er * SB1 + so);
            pey[1] = bld<T>(rpy, oK, (er + 1) * SB1 + so);
            phy[0] = bld<T>(rpy, oK, hr * SB1 + so);
            phy[1] = bld<T>(rpy, oK, (hr + 1) * SB1 + so);
        }
        T pex[2] = {T(0), T(0)}, phx[2] = {T(0), T(0)};
        if (rv && xpml) {
            pex[0] = bld<T>(rpx, oR, er * SB0 + sxo);
            pex[1] = bld<T>(rpx, oR, (er + 1) * SB0 + sxo);
            if (owned) {
                phx[0] = bld<T>(rpx, oR, hr * SB0 + sxo);
                phx[1] = bld<T>(rpx, oR, (hr + 1) * SB0 + sxo);
            }
        }
        const bool wr = own_plane && owned;
        int mm[3] = {0, 0, 0}, npl[3] = {0, 0, 0};
        bool des[3] = {false, false, false};
        T dv[3];
        bool desany = false;
        uint32_t sact = 0u;  // pole slots some lane of the wavefront uses (uniform)
        T qv[NP > 0 ? NP : 1][3], pv[NP > 0 ? NP : 1][3];
        if (wave_any(wr) && !uniform) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {  // lanes of other rows: no poles, no design voxel
                const int m = (wc >> (8 * c)) & 0xff;
                des[c] = wr && m == kDesignMat;
                mm[c] = (wr && !des[c]) ? m : 0;
                npl[c] = (wr && !des[c]) ? slut[mm[c]].np : 0;
            }
            desany = wave_any(des[0] || des[1] || des[2]);
            if (desany) {
                const uint32_t dq = (uint32_t)(((pi - a.dlo[0]) * dey + (j - a.dlo[1])) * dez + (k - a.dlo[2])) * TS;
#pragma unroll
                for (int c = 0; c < 3; ++c) dv[c] = bld<T>(rd, des[c] ? dq : kOOB, 0);
            }
            if (NP > 0) {
                const Rsrc rQ = rsrc((const char*)g.Aq + (long)pi * QPB);
                const Rsrc rP = rsrc((const char*)g.Ap + (long)pi * QPB);
#pragma unroll
                for (int s = 0; s < NP; ++s) {
                    if (!wave_any(s < npl[0] || s < npl[1] || s < npl[2])) continue;
                    sact |= 1u << s;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const bool vq = s < npl[c];
                        const bool vp = vq && ((slut[mm[c]].lorentz_mask >> s) & 1);
                        const uint32_t so = (s * 3 + c) * PB;
                        pv[s][c] = bld<T>(rP, vp ? oR : kOOB, so);
                        qv[s][c] = bld<T>(rQ, vq ? oR : kOOB, so);
                    }
                }
            }
        }

        // ---- phase 1: E-terms of this row, Hbar_total ---------------------------------------
        T t_ex_y = cdt * uc[0], t_ex_z = -cdt * uc[0], t_ey_z = cdt * uc[1], t_ey_x = -cdt * uc[1],
          t_ez_x = cdt * uc[2], t_ez_y = -cdt * uc[2];
        if (ypml) {
            t_ex_y = bEy * t_ex_y + (bEy - T(1)) * pey[0];
            t_ez_y = bEy * t_ez_y + (bEy - T(1)) * pey[1];
        }
        if (rv && zany) {
            t_ex_z = bEz * t_ex_z + (bEz - T(1)) * pzE[0];
            t_ey_z = bEz * t_ey_z + (bEz - T(1)) * pzE[1];
        }
        if (rv && xpml) {
            t_ey_x = bEx * t_ey_x + (bEx - T(1)) * pex[0];
            t_ez_x = bEx * t_ez_x + (bEx - T(1)) * pex[1];
        }
        // x-terms of plane i + 1: (Ey -> Hz), (Ez -> Hy)
        T xn0 = -cdt * un[1], xn1 = cdt * un[2];
        if (rv && xpmln) {
            xn0 = bExn * xn0 + (bExn - T(1)) * bld<T>(rpx, oR, er * SB0 + sxn);
            xn1 = bExn * xn1 + (bExn - T(1)) * bld<T>(rpx, oR, (er + 1) * SB0 + sxn);
        }
        xn0 = (rv && pn >= 0) ? xn0 : T(0);
        xn1 = (rv && pn >= 0) ? xn1 : T(0);
        if (act) {
            sE[0 * B + tid] = t_ex_z;
            sE[1 * B + tid] = t_ey_z;
            sE[2 * B + tid] = t_ex_y;
            sE[3 * B + tid] = t_ez_y;
        }
        if (top) {  // y-terms of the row above the tile
            T ay = cdt * tc[0], az = -cdt * tc[1];
            if (ypmlt) {
                const uint32_t so = (uint32_t)((pi * Ly + syt) * nz) * TS;
                ay = bEyt * ay + (bEyt - T(1)) * bld<T>(rpy, oK, er * SB1 + so);
                az = bEyt * az + (bEyt - T(1)) * bld<T>(rpy, oK, (er + 1) * SB1 + so);
            }
            sEt[k] = ay;
            sEt[nz + k] = az;
        }
        lds_barrier();
        T tot[3];
        {
            const T exz_k = kp1 >= 0 ? sE[0 * B + r * nz + kp1] : T(0);
            const T eyz_k = kp1 >= 0 ? sE[1 * B + r * nz + kp1] : T(0);
            const T exy_up = !act ? T(0) : r + 1 < RT ? sE[2 * B + tid + nz] : sEt[k];
            const T ezy_up = !act ? T(0) : r + 1 < RT ? sE[3 * B + tid + nz] : sEt[nz + k];
            // Hx: (Ey, z) and (Ez, y); Hy: (Ex, z) and (Ez, x); Hz: (Ex, y) and (Ey, x)
            tot[0] = rv ? hb[0] + (t_ey_z - eyz_k) + (t_ez_y - ezy_up) : T(0);
            tot[1] = rv ? hb[1] + (t_ex_z - exz_k) + (t_ez_x - xn1) : T(0);
            tot[2] = rv ? hb[2] + (t_ex_y - exy_up) + (t_ey_x - xn0) : T(0);
        }
        // owned row: Hbar at the old time and the new H auxiliary bars
        if (wr) {
            bst(rB, oR, 3 * PB, tot[0]);
            bst(rB, oR, 4 * PB, tot[1]);
            bst(rB, oR, 5 * PB, tot[2]);
            if (ypml) {  // (Hx, y, +, slot 0), (Hz, y, -, slot 1)
                const uint32_t so = (uint32_t)((pi * Ly + sy) * nz) * TS;
                bst(rpy, oK, hw * SB1 + so, bHy * (-cdt * tot[0] + phy[0]));
                bst(rpy, oK, (hw + 1) * SB1 + so, bHy * (cdt * tot[2] + phy[1]));
            }
            if (zany) {  // (Hx, z, -, slot 0), (Hy, z, +, slot 1)
                const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
                bst(rpz, oZ, hw * SB2 + so, bHz * (cdt * tot[0] + phz[0]));
                bst(rpz, oZ, (hw + 1) * SB2 + so, bHz * (-cdt * tot[1] + phz[1]));
            }
            if (xpml) {  // (Hy, x, -, slot 0), (Hz, x, +, slot 1)
                bst(rpx, oR, hw * SB0 + sxo, bHx * (cdt * tot[1] + phx[0]));
                bst(rpx, oR, (hw + 1) * SB0 + sxo, bHx * (-cdt * tot[2] + phx[1]));
            }
        }

        // ---- phase 2: H-terms of this row, Ebar of the owned rows --------------------------
        T d_hx_y = -cdt * tot[0], d_hx_z = cdt * tot[0], d_hy_z = -cdt * tot[1], d_hy_x = cdt * tot[1],
          d_hz_x = -cdt * tot[2], d_hz_y = cdt * tot[2];
        if (ypml) {
            d_hx_y = bHy * d_hx_y + (bHy - T(1)) * phy[0];
            d_hz_y = bHy * d_hz_y + (bHy - T(1)) * phy[1];
        }
        if (owned && zany) {
            d_hx_z = bHz * d_hx_z + (bHz - T(1)) * phz[0];
            d_hy_z = bHz * d_hy_z + (bHz - T(1)) * phz[1];
        }
        if (owned && xpml) {
            d_hy_x = bHx * d_hy_x + (bHx - T(1)) * phx[0];
            d_hz_x = bHx * d_hz_x + (bHx - T(1)) * phx[1];
        }
        d_hx_y = rv ? d_hx_y : T(0);
        d_hz_y = rv ? d_hz_y : T(0);
        if (act) {
            sD[0 * B + tid] = d_hy_z;
            sD[1 * B + tid] = d_hx_z;
            sD[2 * B + tid] = d_hx_y;
            sD[3 * B + tid] = d_hz_y;
        }
        lds_barrier();
        if (wr) {
            const uint32_t o = oR;
            const T hyz_k = km1 >= 0 ? sD[0 * B + r * nz + km1] : T(0);
            const T hxz_k = km1 >= 0 ? sD[1 * B + r * nz + km1] : T(0);
            const T hxy_dn = sD[2 * B + tid - nz];  // y-terms of the row below (r >= 1)
            const T hzy_dn = sD[3 * B + tid - nz];
            T inv[3], epsc[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                inv[c] = slut[mm[c]].inv_eps;
                epsc[c] = slut[mm[c]].eps_inf;
            }
            if (desany) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    inv[c] = des[c] ? dv[c] : inv[c];
                    epsc[c] = des[c] ? T(1) / dv[c] : epsc[c];
                }
            }
            // Ebar' = eps Ubar' plus the differences
            // Ex: (Hy, z) and (Hz, y); Ey: (Hx, z) and (Hz, x); Ez: (Hx, y) and (Hy, x)
            T out[3];
            out[0] = epsc[0] * uc[0] + (hyz_k - d_hy_z) + (hzy_dn - d_hz_y);
            out[1] = epsc[1] * uc[1] + (hxz_k - d_hx_z) + (dbx_prev[0] - d_hz_x);
            out[2] = epsc[2] * uc[2] + (hxy_dn - d_hx_y) + (dbx_prev[1] - d_hy_x);
            if (NP > 0 && !uniform) {
                const Rsrc rQ = rsrc((const char*)g.Aq + (long)pi * QPB);
                const Rsrc rP = rsrc((const char*)g.Ap + (long)pi * QPB);
#pragma unroll
                for (int s = 0; s < NP; ++s) {
                    if (!((sact >> s) & 1u)) continue;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const MatLut<T>& L = slut[mm[c]];
                        const bool vq = s < npl[c];
                        const bool vp = vq && ((L.lorentz_mask >> s) & 1);
                        const uint32_t so = (s * 3 + c) * PB;
                        const T pb = pv[s][c];
                        const T qb = vq ? qv[s][c] + pb - uc[c] : T(0);
                        bst(rQ, vq ? o : kOOB, so, L.alpha[s] * qb);
                        bst(rP, vp ? o : kOOB, so, pb + L.beta[s] * qb);
                        out[c] += L.c3[s] * qb;
                    }
                }
            }
            // new E auxiliary bars: E-terms of this row with their signs
            if (ypml) {  // (Ex, y, +, slot 0), (Ez, y, -, slot 1)
                const uint32_t so = (uint32_t)((pi * Ly + sy) * nz) * TS;
                bst(rpy, oK, ew * SB1 + so, bEy * (cdt * uc[0] + pey[0]));
                bst(rpy, oK, (ew + 1) * SB1 + so, bEy * (-cdt * uc[2] + pey[1]));
            }
            if (zany) {  // (Ex, z, -, slot 0), (Ey, z, +, slot 1)
                const uint32_t so = (uint32_t)((pi * ny + j) * Lz) * TS;
                bst(rpz, oZ, ew * SB2 + so, bEz * (-cdt * uc[0] + pzE[0]));
                bst(rpz, oZ, (ew + 1) * SB2 + so, bEz * (cdt * uc[1] + pzE[1]));
            }
            if (xpml) {  // (Ey, x, -, slot 0), (Ez, x, +, slot 1)
                bst(rpx, o, ew * SB0 + sxo, bEx * (-cdt * uc[1] + pex[0]));
                bst(rpx, o, (ew + 1) * SB0 + sxo, bEx * (cdt * uc[2] + pex[1]));
            }
            // the state is Ubar = Ebar / eps
            bst(rB, o, 0, inv[0] * out[0]);
            bst(rB, o, PB, inv[1] * out[1]);
            bst(rB, o, 2 * PB, inv[2] * out[2]);
        }
        dbx_prev[0] = owned ? d_hz_x : T(0);
        dbx_prev[1] = owned ? d_hy_x : T(0);
#pragma unroll
        for (int c = 0; c < 3; ++c) uc[c] = un[c];
        tc[0] = tn[0];
        tc[1] = tn[1];
        wc = wn;
    }
}

// Fused reverse step over one engine, row-of-threads form (tile t after the XCD-aware remap).
template <typename T, int NP, int MAXT>
__global__ void __launch_bounds__(MAXT) k_adj_rows(const AdjArgs<T> g) {
    const int w = blockIdx.x, nw = gridDim.x;
    const int t = (nw % 8 == 0) ? ((w & 7) * (nw >> 3) + (w >> 3)) : w;
    adj_rows<T, NP>(g, t);
}

// Fused reverse step over one engine: workgroup w handles tile w after the XCD-aware remap.
// With the loads of a plane issued together, the FP32 reverse step needs 95 registers with three
// pole slots (5 wavefronts per SIMD) and 102 with four (4 wavefronts), without spilling.
template <typename T>
constexpr int adj_waves() { return sizeof(T) == 4 ? 4 : 1; }

template <typename T, int NP, int MAXT, bool NTS = false>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(adj_waves<T>())))
k_adj_step(const AdjArgs<T> g) {
    const int w = blockIdx.x, nw = gridDim.x;
    const int t = (nw % 8 == 0) ? ((w & 7) * (nw >> 3) + (w >> 3)) : w;
    adj_tile<T, NP, NTS>(g, t);
}

// record E of the design box and of the objective cell into slot `slot` of a ring buffer
template <typename T>
__global__ void __launch_bounds__(256) k_adj_record(const T* E, long nyz, int ny, int nz, int dlo0, int dlo1,
                                                    int dlo2, int de0, int de1, int de2, long gap,
                                                    T* ring, long slot) {
    const long nd = (long)de0 * de1 * de2;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    T* out = ring + slot * (3 * nd + 3);
    if (v < nd) {
        const int kk = (int)(v % de2);
        const long r = v / de2;
        const int jj = (int)(r % de1), ii = (int)(r / de1);
        const long q = ((long)(dlo0 + ii) * ny + (dlo1 + jj)) * nz + (dlo2 + kk);
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c * nd + v] = E[field_index(q, c, nyz)];
    } else if (v < nd + 3) {
        const int c = (int)(v - nd);
        out[3 * nd + c] = E[field_index(gap, c, nyz)];
    }
}

// objective J += dt * sum_c E_c(gap)^2 for the state after a step (FP64)
template <typename T>
__global__ void k_adj_objective(const T* E, long nyz, long gap, double dt, double* J) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int c = 0; c < 3; ++c) {
            const double v = (double)E[field_index(gap, c, nyz)];
            s += v * v;
        }
        *J += s * dt;
    }
}

// gradient term of step n:
//   gacc[v] += sum_c Ubar_c^{n+1}(v) (E_c^{n+1}(v) - E_c^n(v))   (before the adjoint step)
// With Ubar = Ebar / eps_v this is (1/eps_v) sum_c Ebar_c^{n+1} (E^{n+1} - E^n): the gradient
// dJ/d eps_v = -(1/eps_v) sum_n sum_c Ebar (E^{n+1} - E^n) is minus the accumulated sum.
// Ebar: the adjoint field set holding Ubar^{n+1} ([x][comp][y][z]).
template <typename T>
__global__ void __launch_bounds__(256) k_adj_grad(const StepArgs<T> a, const T* Ebar, const T* ring, long s_next,
                                                  long s_cur, double* gacc) {
    const long nd = (long)(a.dhi[0] - a.dlo[0]) * (a.dhi[1] - a.dlo[1]) * (a.dhi[2] - a.dlo[2]);
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd) return;
    const int de1 = a.dhi[1] - a.dlo[1], de2 = a.dhi[2] - a.dlo[2];
    const int kk = (int)(v % de2);
    const long r = v / de2;
    const int jj = (int)(r % de1), ii = (int)(r / de1);
    const long q = ((long)(a.dlo[0] + ii) * a.ny + (a.dlo[1] + jj)) * a.nz + (a.dlo[2] + kk);
    const T* e1 = ring + s_next * (3 * nd + 3);
    const T* e0 = ring + s_cur * (3 * nd + 3);
    double acc = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        acc += (double)Ebar[field_index(q, c, (long)a.ny * a.nz)] * ((double)e1[c * nd + v] - (double)e0[c * nd + v]);
    gacc[v] += acc;
}

// Ebar^n(gap) += dJ/dE^n = 2 dt E^n(gap), on the state Ubar = Ebar / eps: Ubar(gap) gains
// 2 dt E^n(gap) / eps(gap) per component (ie = 1/eps at the gap's three E points)
template <typename T>
__global__ void k_adj_inject(T* Ubar, long nyz, const T* ring, long nd, long slot, long gap, double dt, double ie0,
                             double ie1, double ie2) {
    const int c = threadIdx.x;
    if (blockIdx.x == 0 && c < 3) {
        const double e = (double)ring[slot * (3 * nd + 3) + 3 * nd + c];
        const double ie = c == 0 ? ie0 : (c == 1 ? ie1 : ie2);
        const long o = field_index(gap, c, nyz);
        Ubar[o] = (T)((double)Ubar[o] + 2.0 * dt * e * ie);
    }
}

// ---- adjoint over a batch: the kernels above, one launch for every engine of the set --------
// Per-engine table of the batched record, gradient and injection kernels (blockIdx.y is the
// engine); the fused reverse step maps workgroups to engines as k_step_rows does.
template <typename T>
struct AdjBatchDev {
    const T* F[2];   // forward field sets 0 / 1
    T* A[2];         // adjoint field sets (Ubar, Hbar)
    T* ring;         // record: per step E of the design box (3 nd) and of the objective cell (3)
    double* gacc;    // nd gradient sums
    double* J;       // objective
    long nyz;
    int ny, nz;
    int dlo[3], de[3];
    long nd;
    long gap;        // objective cell
    double dt;
    double ie[3];    // 1/eps at the objective cell's three E points
};

// k_adj_record for every engine (E = field set `par`), and with `objective` set also
// k_adj_objective: the thread of the objective cell's first component adds dt sum_c E_c(gap)^2
template <typename T>
__global__ void __launch_bounds__(256) k_adj_record_batch(const AdjBatchDev<T>* __restrict__ tab, int par,
                                                          long slot, int objective) {
    const AdjBatchDev<T>& a = tab[blockIdx.y];
    const long nd = a.nd;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd + 3) return;
    const T* E = a.F[par];
    T* out = a.ring + slot * (3 * nd + 3);
    if (v < nd) {
        const int kk = (int)(v % a.de[2]);
        const long r = v / a.de[2];
        const int jj = (int)(r % a.de[1]), ii = (int)(r / a.de[1]);
        const long q = ((long)(a.dlo[0] + ii) * a.ny + (a.dlo[1] + jj)) * a.nz + (a.dlo[2] + kk);
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c * nd + v] = E[field_index(q, c, a.nyz)];
    } else {
        const int c = (int)(v - nd);
        out[3 * nd + c] = E[field_index(a.gap, c, a.nyz)];
        if (objective && c == 0) {
            double sum = 0;
            for (int k = 0; k < 3; ++k) {
                const double x = (double)E[field_index(a.gap, k, a.nyz)];
                sum += x * x;
            }
            *a.J += sum * a.dt;
        }
    }
}

// k_adj_grad for every engine (Ubar^{n+1} in adjoint set `rd`)
template <typename T>
__global__ void __launch_bounds__(256) k_adj_grad_batch(const AdjBatchDev<T>* __restrict__ tab, int rd,
                                                        long s_next, long s_cur) {
    const AdjBatchDev<T>& a = tab[blockIdx.y];
    const long nd = a.nd;
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nd) return;
    const int kk = (int)(v % a.de[2]);
    const long r = v / a.de[2];
    const int jj = (int)(r % a.de[1]), ii = (int)(r / a.de[1]);
    const long q = ((long)(a.dlo[0] + ii) * a.ny + (a.dlo[1] + jj)) * a.nz + (a.dlo[2] + kk);
    const T* Ebar = a.A[rd];
    const T* e1 = a.ring + s_next * (3 * nd + 3);
    const T* e0 = a.ring + s_cur * (3 * nd + 3);
    double acc = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
        acc += (double)Ebar[field_index(q, c, a.nyz)] * ((double)e1[c * nd + v] - (double)e0[c * nd + v]);
    a.gacc[v] += acc;
}

// k_adj_inject for every engine (into adjoint set `rd`, record slot `slot`)
template <typename T>
__global__ void k_adj_inject_batch(const AdjBatchDev<T>* __restrict__ tab, int rd, long slot) {
    const AdjBatchDev<T>& a = tab[blockIdx.y];
    const int c = threadIdx.x;
    if (blockIdx.x == 0 && c < 3) {
        const double e = (double)a.ring[slot * (3 * a.nd + 3) + 3 * a.nd + c];
        const long o = field_index(a.gap, c, a.nyz);
        T* U = a.A[rd];
        U[o] = (T)((double)U[o] + 2.0 * a.dt * e * a.ie[c]);
    }
}

// Fused reverse step over every engine of a batch (tile t of engine b after the XCD-aware remap)
template <typename T, int NP, int MAXT, bool NTS = false>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(adj_waves<T>())))
k_adj_step_batch(const AdjArgs<T>* __restrict__ args, const int* __restrict__ tile_first, int ninst) {
    int t;
    const int b = tile_instance(tile_first, ninst, t);
    adj_tile<T, NP, NTS>(args[b], t);
}

// The same, row-of-threads form
template <typename T, int NP, int MAXT>
__global__ void __launch_bounds__(MAXT) k_adj_rows_batch(const AdjArgs<T>* __restrict__ args,
                                                         const int* __restrict__ tile_first, int ninst) {
    int t;
    const int b = tile_instance(tile_first, ninst, t);
    adj_rows<T, NP>(args[b], t);
}

// ---- plain streaming copy (achievable-bandwidth probe and counter calibration) -----------
// one float4 per thread over the whole array (the fastest plain copy measured on MI355X)
__global__ void __launch_bounds__(256) k_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                              long n) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[q];
}

}  // namespace sjf
