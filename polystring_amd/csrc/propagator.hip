// propagator.hip -- the RQM4 / OS2 split-step kernels for gfx950.
//
// One split step of the modified diffusion equation (src/RQM4.cc:71-89,
// src/PseudoSpectral.cc:95-102) is
//     q' = e(r) * F^-1[ K(k) * F[ e(r) * q ] ] / M,   e = exp(-w h/2), K = exp(-k^2 h).
// The 3-D transform is split into two passes over HBM:
//   * k_plane : per (lane, x-plane), the yz plane staged in LDS.  Inverse
//               direction: y inverse, complex-to-real along z, times the
//               post factor / M, optional RQM4 Richardson combine and slice
//               stores.  Forward direction: times the pre factor, real-to-
//               complex along z, y forward.  Both directions run back to back
//               in one launch, so the step-to-step "inverse, multiply by
//               e(r) twice, forward" never leaves LDS.
//   * k_xpass : per (lane, tile of C adjacent (y,kz) columns), the x pencils
//               staged in LDS: forward x, times K, inverse x.
// The real-to-complex transform along z uses the half-length complex FFT of
// even/odd samples packed as z_n = x_{2n} + i x_{2n+1}, followed by the
// standard split  X_k = E_k + W^k O_k,  E_k = (Z_k + Z*_{H-k})/2,
// O_k = -i (Z_k - Z*_{H-k})/2; the complex-to-real inverse is its transpose.
#include "engine.hpp"

namespace po {

template <typename T, int LY, int LZ> struct PlaneCfg {
    static constexpr int H = LZ / 2;
    static constexpr int HZ = LZ == 1 ? 1 : H + 1;
    static constexpr int NE = LY * LZ;  // reals per plane
    static constexpr int NC = LY * HZ;  // complex per spectrum plane
    // largest of 512/256/128/64 threads giving every thread >= 16 reals;
    // 512 leaves a 256-VGPR budget for the staged butterflies.
    static constexpr int NT = NE >= 16 * 512 ? 512 : NE >= 16 * 256 ? 256 : NE >= 16 * 128 ? 128 : 64;
    static constexpr int EPT = (NE + NT - 1) / NT;
};

// y transform of the HZ columns of a plane.  With LZ > 1 the H = LZ/2
// regular columns form one batch and the Nyquist column (kz = H) a second
// one, so the main batch maps evenly onto the threads and needs fewer staged
// butterflies (registers) per thread.
template <typename T, int LY, int LZ, bool INV, int NT>
__device__ __forceinline__ void fft_columns(cx<T> *buf, const Roots<T> &rt) {
    constexpr int H = LZ / 2;
    constexpr int HZ = LZ == 1 ? 1 : H + 1;
    if constexpr (LZ == 1) {
        fft_lds<T, LY, INV, NT, 1, true>(buf, 1, 1, rt.y);
    } else {
        fft_lds<T, LY, INV, NT, H, true>(buf, 1, HZ, rt.y);
        fft_lds<T, LY, INV, NT, 1, true>(buf + H, 1, HZ, rt.y);
    }
}

// real-to-complex along z (rows), then forward y (columns)
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_forward(cx<T> *buf, const Roots<T> &rt) {
    constexpr int H = LZ / 2;
    constexpr int HZ = LZ == 1 ? 1 : H + 1;
    if constexpr (LZ > 1) {
        fft_lds<T, H, false, NT, LY, false>(buf, HZ, 1, rt.zh);
        constexpr int NPAIR = H / 2 + 1;
        for (int t = threadIdx.x; t < LY * NPAIR; t += NT) {
            const int y = t / NPAIR, k = t % NPAIR;
            cx<T> *row = buf + y * HZ;
            if (k == 0) {
                const cx<T> z0 = row[0];
                row[0] = mk<T>(z0.x + z0.y, (T)0);
                row[H] = mk<T>(z0.x - z0.y, (T)0);
            } else {
                const int k2 = H - k;
                const cx<T> a = row[k], b = row[k2];
                const T hf = (T)0.5;
                // X_k from (a, b), X_k2 from (b, a)
                const cx<T> ea = scl<T>(add<T>(a, conj<T>(b)), hf);
                const cx<T> oa = rot<false, T>(scl<T>(sub<T>(a, conj<T>(b)), hf));
                const cx<T> eb = scl<T>(add<T>(b, conj<T>(a)), hf);
                const cx<T> ob = rot<false, T>(scl<T>(sub<T>(b, conj<T>(a)), hf));
                const cx<T> xa = add<T>(ea, cmul<false, T>(oa, rt.z[k]));
                const cx<T> xb = add<T>(eb, cmul<false, T>(ob, rt.z[k2]));
                row[k] = xa;
                row[k2] = xb;
            }
        }
        __syncthreads();
    }
    if constexpr (LY > 1) fft_columns<T, LY, LZ, false, NT>(buf, rt);
}

// inverse y (columns), then complex-to-real along z (rows); unnormalised
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_inverse(cx<T> *buf, const Roots<T> &rt) {
    constexpr int H = LZ / 2;
    constexpr int HZ = LZ == 1 ? 1 : H + 1;
    if constexpr (LY > 1) fft_columns<T, LY, LZ, true, NT>(buf, rt);
    if constexpr (LZ > 1) {
        constexpr int NPAIR = H / 2 + 1;
        for (int t = threadIdx.x; t < LY * NPAIR; t += NT) {
            const int y = t / NPAIR, k = t % NPAIR;
            cx<T> *row = buf + y * HZ;
            if (k == 0) {
                const cx<T> x0 = row[0], xh = row[H];
                const cx<T> e = add<T>(x0, conj<T>(xh));
                const cx<T> o = sub<T>(x0, conj<T>(xh));
                row[0] = add<T>(e, rot<true, T>(o));
            } else {
                const int k2 = H - k;
                const cx<T> a = row[k], b = row[k2];
                const cx<T> ea = add<T>(a, conj<T>(b));
                const cx<T> oa = cmul<true, T>(sub<T>(a, conj<T>(b)), rt.z[k]);
                const cx<T> eb = add<T>(b, conj<T>(a));
                const cx<T> ob = cmul<true, T>(sub<T>(b, conj<T>(a)), rt.z[k2]);
                row[k] = add<T>(ea, rot<true, T>(oa));
                row[k2] = add<T>(eb, rot<true, T>(ob));
            }
        }
        __syncthreads();
        fft_lds<T, H, true, NT, LY, false>(buf, HZ, 1, rt.zh);
    }
}

template <typename T, int LY, int LZ>
__global__ void __launch_bounds__((PlaneCfg<T, LY, LZ>::NT))
k_plane(const PlaneDesc<T> *__restrict__ descs, Roots<T> rt, T invM) {
    using C = PlaneCfg<T, LY, LZ>;
    constexpr int NT = C::NT, NE = C::NE, NC = C::NC, EPT = C::EPT, HZ = C::HZ;
    __shared__ cx<T> buf[NC];
    T *rbuf = reinterpret_cast<T *>(buf);
    const PlaneDesc<T> d = descs[blockIdx.y];
    const long ro = (long)blockIdx.x * NE;
    const long co = (long)blockIdx.x * NC;
    const int tid = threadIdx.x;
    auto lpos = [](int e) { return (e / LZ) * 2 * HZ + (e % LZ); };
    T q[EPT];

    if (d.in0) {
#pragma unroll 8
        for (int c = tid; c < NC; c += NT) buf[c] = d.in0[co + c];
        __syncthreads();
        plane_inverse<T, LY, LZ, NT>(buf, rt);
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const T f = d.post0 ? d.post0[ro + e] : (T)1;
                q[i] = (rbuf[lpos(e)] * f) * invM;
            }
        }
        if (d.in1) {
            __syncthreads();
#pragma unroll 8
            for (int c = tid; c < NC; c += NT) buf[c] = d.in1[co + c];
            __syncthreads();
            plane_inverse<T, LY, LZ, NT>(buf, rt);
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                const int e = tid + i * NT;
                if (NE % NT == 0 || e < NE) {
                    const T f = d.post1 ? d.post1[ro + e] : (T)1;
                    const T qh = (rbuf[lpos(e)] * f) * invM;
                    q[i] = ((T)4 * qh - q[i]) / (T)3;  // src/RQM4.cc:91
                }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) q[i] = d.qin[ro + e];
        }
    }
    if (d.qout) {
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                d.qout[ro + e] = q[i];
                if (d.qout2) d.qout2[ro + e] = q[i];
            }
        }
    }
    if (d.out0) {
        // Every thread overwrites only the LDS reals it read itself, so no
        // barrier is needed between the inverse read-out and this write.
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const T v = d.pre0 ? d.pre0[ro + e] * q[i] : q[i];
                if (LZ == 1) buf[e] = mk<T>(v, (T)0);
                else rbuf[lpos(e)] = v;
            }
        }
        __syncthreads();
        plane_forward<T, LY, LZ, NT>(buf, rt);
#pragma unroll 8
        for (int c = tid; c < NC; c += NT) d.out0[co + c] = buf[c];
    }
    if (d.out1) {
        // The slice is re-read from where this thread stored it (qout, or the
        // input slice) instead of being held in registers across out0.
        const T *qs = d.qout ? d.qout : d.qin;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const T v = d.pre1 ? d.pre1[ro + e] * qs[ro + e] : qs[ro + e];
                if (LZ == 1) buf[e] = mk<T>(v, (T)0);
                else rbuf[lpos(e)] = v;
            }
        }
        __syncthreads();
        plane_forward<T, LY, LZ, NT>(buf, rt);
#pragma unroll 8
        for (int c = tid; c < NC; c += NT) d.out1[co + c] = buf[c];
    }
}

template <typename T, int LX> struct XCfg {
    static constexpr int C = 16;  // adjacent (y,kz) columns per workgroup
    static constexpr int NT0 = ((LX * C / 8 + 63) / 64) * 64;
    static constexpr int NT = NT0 < 64 ? 64 : (NT0 > 1024 ? 1024 : NT0);
};

template <typename T, int LX>
__global__ void __launch_bounds__((XCfg<T, LX>::NT))
k_xpass(const XDesc<T> *__restrict__ descs, Roots<T> rt, int ncol, int hz) {
    constexpr int C = XCfg<T, LX>::C, NT = XCfg<T, LX>::NT;
    __shared__ cx<T> buf[LX * C];
    const XDesc<T> d = descs[blockIdx.y];
    const int c0 = blockIdx.x * C;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < LX * C; idx += NT) {
        const int x = idx / C, col = c0 + idx % C;
        buf[idx] = col < ncol ? d.S[(long)x * ncol + col] : mk<T>((T)0, (T)0);
    }
    __syncthreads();
    fft_lds<T, LX, false, NT, C, true>(buf, 1, C, rt.x);
    for (int idx = tid; idx < LX * C; idx += NT) {
        const int x = idx / C, col = c0 + idx % C;
        if (col < ncol) {
            const T f = d.tab ? d.tab[(long)x * ncol + col] : d.fx[x] * d.fy[col / hz] * d.fz[col % hz];
            buf[idx] = scl<T>(buf[idx], f);
        }
    }
    __syncthreads();
    fft_lds<T, LX, true, NT, C, true>(buf, 1, C, rt.x);
    for (int idx = tid; idx < LX * C; idx += NT) {
        const int x = idx / C, col = c0 + idx % C;
        if (col < ncol) d.S[(long)x * ncol + col] = buf[idx];
    }
}

// ------------------------------------------------------------- dispatch --
#define PO_PLANE_SHAPES(X) \
    X(1, 1) X(8, 1) X(16, 1) X(32, 1) X(48, 1) X(64, 1) X(96, 1) X(128, 1) X(256, 1) \
    X(8, 8) X(16, 16) X(32, 32) X(48, 48) X(64, 64) X(96, 96) X(128, 128)
#define PO_X_SIZES(X) X(8) X(16) X(32) X(48) X(64) X(96) X(128) X(256)

bool plane_supported(int Ly, int Lz) {
#define PO_CASE(a, b) if (Ly == a && Lz == b) return true;
    PO_PLANE_SHAPES(PO_CASE)
#undef PO_CASE
    return false;
}

bool xpass_supported(int Lx) {
#define PO_CASE(a) if (Lx == a) return true;
    PO_X_SIZES(PO_CASE)
#undef PO_CASE
    return false;
}

template <typename T>
hipError_t launch_plane(const Shape &sh, const Roots<T> &rt, const PlaneDesc<T> *d_desc, int nlanes,
                        hipStream_t st) {
    const T invM = (T)(1.0 / (double)sh.M);
#define PO_CASE(a, b)                                                                           \
    if (sh.Ly == a && sh.Lz == b) {                                                             \
        hipLaunchKernelGGL((k_plane<T, a, b>), dim3(sh.Lx, nlanes), dim3(PlaneCfg<T, a, b>::NT), \
                           0, st, d_desc, rt, invM);                                            \
        return hipGetLastError();                                                               \
    }
    PO_PLANE_SHAPES(PO_CASE)
#undef PO_CASE
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_xpass(const Shape &sh, const Roots<T> &rt, const XDesc<T> *d_desc, int nlanes,
                        hipStream_t st) {
    const int ncol = sh.Ly * sh.Hz;
#define PO_CASE(a)                                                                         \
    if (sh.Lx == a) {                                                                      \
        const int nt = (ncol + XCfg<T, a>::C - 1) / XCfg<T, a>::C;                         \
        hipLaunchKernelGGL((k_xpass<T, a>), dim3(nt, nlanes), dim3(XCfg<T, a>::NT), 0, st, \
                           d_desc, rt, ncol, sh.Hz);                                       \
        return hipGetLastError();                                                          \
    }
    PO_X_SIZES(PO_CASE)
#undef PO_CASE
    return hipErrorInvalidValue;
}

template hipError_t launch_plane<double>(const Shape &, const Roots<double> &, const PlaneDesc<double> *, int,
                                         hipStream_t);
template hipError_t launch_plane<float>(const Shape &, const Roots<float> &, const PlaneDesc<float> *, int,
                                        hipStream_t);
template hipError_t launch_xpass<double>(const Shape &, const Roots<double> &, const XDesc<double> *, int,
                                         hipStream_t);
template hipError_t launch_xpass<float>(const Shape &, const Roots<float> &, const XDesc<float> *, int,
                                        hipStream_t);

}  // namespace po
