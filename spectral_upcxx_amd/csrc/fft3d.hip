// fft3d.hip -- batched 1D FFT passes and the 3D transform (Scalar::transform,
// src/Scalar.cpp:400-646; FFT1D, src/FFT1D.cpp:14-101).
//
// With the whole field in one GPU's HBM the reference's slab transposes
// (upcxx::rput all-to-all, src/Scalar.cpp:446-626) are unnecessary: the y and
// z passes read their lines in place through column tiles.
//   * contiguous pass (x): a block owns LPB whole lines; lane order is
//     element-fastest, so every global access is a unit-stride run.
//   * strided pass (y, z): a block owns LPB adjacent columns of one plane;
//     lane order is column-fastest, so each access instruction reads
//     LPB*sizeof(complex) contiguous bytes per row; the FFT runs across the
//     rows through LDS laid out [position][column].
// Each pass reads and writes the field once (in place), so a 3D transform
// moves 3 x (read + write) x N^3 elements -- the algorithmic minimum for
// three axis passes.  The optional epilogue (normalisation, Nyquist zeroing)
// rides on the last pass.
#include <map>
#include <vector>

#include "fft_core.hpp"

namespace sh {

template <typename T, class P>
__device__ __forceinline__ C2<T> epilogue(C2<T> v, int pos, bool zline, const Epi<T>& ep) {
    if (ep.flags & 1) v = rmul<T>(ep.scale, v);
    if ((ep.flags & 2) && (zline || pos == P::N / 2)) v = mk<T>(T(0), T(0));
    return v;
}

template <int N>
struct ContigLds {
    static constexpr int LS = N + N / 16 + 1;
};

// NT: nontemporal stores; LNT: nontemporal loads
template <typename T, class P, int LPB, int SIGN, bool NT = false, bool LNT = false>
__global__ __launch_bounds__(P::TPL* LPB) void k_fft_contig(C2<T>* __restrict__ data,
                                                            const C2<T>* __restrict__ tw,
                                                            long nlines, long ldist, long fstride,
                                                            Epi<T> ep) {
    constexpr int LS = ContigLds<P::N>::LS;
    __shared__ T sre[LPB * LS];
    __shared__ T sim[LPB * LS];
    const int tid = threadIdx.x % P::TPL;
    const int l = threadIdx.x / P::TPL;
    const long line = (long)blockIdx.x * LPB + l;
    const bool valid = line < nlines;
    C2<T>* base = data + (long)blockIdx.y * fstride + (valid ? line : 0) * ldist;
    C2<T> x[P::E];
    for_in<P>(tid, [&](int slot, int pos) {
        if constexpr (LNT) {
            typedef T t2 __attribute__((ext_vector_type(2)));
            const t2 w = __builtin_nontemporal_load(reinterpret_cast<const t2*>(base + pos));
            x[slot] = mk<T>(w.x, w.y);
        } else {
            x[slot] = base[pos];
        }
    });
    auto idx = [&](int pos) { return l * LS + pos + (pos >> 4); };
    line_fft<P, SIGN, T>(x, tid, sre, sim, idx, tw);
    if (!valid) return;
    bool zline = false;
    if (ep.flags & 2) {
        const long a = line % ep.a_n, b = line / ep.a_n;
        zline = (a == ep.za) || (b == ep.zb);
    }
    for_out<P>(tid, [&](int slot, int pos) { st_line<NT, T>(base + pos, epilogue<T, P>(x[slot], pos, zline, ep)); });
}

// NT: nontemporal stores; LNT: nontemporal loads (each value is read once, so
// the loads need not displace in the last-level cache what was just written)
template <typename T, class P, int LPB, int SIGN, bool NT = false, bool LNT = false>
__global__ __launch_bounds__(P::TPL* LPB) void k_fft_strided(C2<T>* __restrict__ data,
                                                             const C2<T>* __restrict__ tw,
                                                             int ncols, long cdist, int nouter,
                                                             long odist, long es, long fstride,
                                                             Epi<T> ep) {
    constexpr int CS = LPB + 1;
    __shared__ T sre[P::N * CS];
    __shared__ T sim[P::N * CS];
    const int c = threadIdx.x % LPB;
    const int tid = threadIdx.x / LPB;
    const int tiles = (ncols + LPB - 1) / LPB;
    const int o = blockIdx.x / tiles;
    const int col = (blockIdx.x % tiles) * LPB + c;
    const bool valid = col < ncols && o < nouter;
    C2<T>* base = data + (long)blockIdx.y * fstride + (valid ? (col * cdist + (long)o * odist) : 0);
    C2<T> x[P::E];
    for_in<P>(tid, [&](int slot, int pos) {
        if constexpr (LNT) {
            typedef T t2 __attribute__((ext_vector_type(2)));
            const t2 w = __builtin_nontemporal_load(reinterpret_cast<const t2*>(base + pos * es));
            x[slot] = mk<T>(w.x, w.y);
        } else {
            x[slot] = base[pos * es];
        }
    });
    auto idx = [&](int pos) { return pos * CS + c; };
    line_fft<P, SIGN, T>(x, tid, sre, sim, idx, tw);
    if (!valid) return;
    const bool zline = (ep.flags & 2) && (col == ep.za || o == ep.zb);
    for_out<P>(tid, [&](int slot, int pos) { st_line<NT, T>(base + pos * es, epilogue<T, P>(x[slot], pos, zline, ep)); });
}

// ---------------------------------------------------------------- launchers
template <typename T, class P, int LPB, bool NT = false, bool LNT = false>
static int launch_contig(C2<T>* data, long nlines, long ldist, int count, long fstride, int dir,
                         const Epi<T>& ep, hipStream_t st) {
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return SH_ENOMEM;
    }
    dim3 grid((unsigned)((nlines + LPB - 1) / LPB), (unsigned)count);
    dim3 block(P::TPL * LPB);
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_contig<T, P, LPB, -1, NT, LNT>), grid, block, 0, st, data, tw, nlines, ldist,
                           fstride, ep);
    else
        hipLaunchKernelGGL((k_fft_contig<T, P, LPB, 1, NT, LNT>), grid, block, 0, st, data, tw, nlines, ldist,
                           fstride, ep);
    SH_LAUNCH_CHECK("k_fft_contig");
    return SH_OK;
}

template <typename T, class P, int LPB, bool NT = false, bool LNT = false>
static int launch_strided(C2<T>* data, int ncols, long cdist, int nouter, long odist, long es,
                          int count, long fstride, int dir, const Epi<T>& ep, hipStream_t st) {
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return SH_ENOMEM;
    }
    const long tiles = (ncols + LPB - 1) / LPB;
    dim3 grid((unsigned)(tiles * nouter), (unsigned)count);
    dim3 block(P::TPL * LPB);
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_strided<T, P, LPB, -1, NT, LNT>), grid, block, 0, st, data, tw, ncols, cdist,
                           nouter, odist, es, fstride, ep);
    else
        hipLaunchKernelGGL((k_fft_strided<T, P, LPB, 1, NT, LNT>), grid, block, 0, st, data, tw, ncols, cdist,
                           nouter, odist, es, fstride, ep);
    SH_LAUNCH_CHECK("k_fft_strided");
    return SH_OK;
}

template <typename T>
using ContigLaunch = int (*)(C2<T>*, long, long, int, long, int, const Epi<T>&, hipStream_t);
template <typename T>
using StridedLaunch = int (*)(C2<T>*, int, long, int, long, long, int, long, int, const Epi<T>&,
                              hipStream_t);

template <typename F>
struct VariantList {
    std::vector<const char*> names;
    std::vector<F> fns;
};

// the tunable schedules of a length (empty when the length has only its default)
template <typename T>
const VariantList<ContigLaunch<T>>& contig_variants(int n) {
    static const std::map<int, VariantList<ContigLaunch<T>>> tab = [] {
        std::map<int, VariantList<ContigLaunch<T>>> m;
#define SH_V(n_, P_, L_)                          \
    m[n_].names.push_back(#P_ " lines=" #L_);     \
    m[n_].fns.push_back(&launch_contig<T, P_, L_>);
        SH_CONTIG_VARIANTS(SH_V)
#undef SH_V
#define SH_VH(n_, P_, L_, NT_, LNT_)                                                                 \
    m[n_].names.push_back(#P_ " lines=" #L_ " nontemporal stores=" #NT_ ", nontemporal loads=" #LNT_); \
    m[n_].fns.push_back(&launch_contig<T, P_, L_, NT_, LNT_>);
        SH_CONTIG_HINT_VARIANTS(SH_VH)
#undef SH_VH
        return m;
    }();
    static const VariantList<ContigLaunch<T>> none;
    auto it = tab.find(n);
    return it == tab.end() ? none : it->second;
}

template <typename T>
const VariantList<StridedLaunch<T>>& strided_variants(int n) {
    static const std::map<int, VariantList<StridedLaunch<T>>> tab = [] {
        std::map<int, VariantList<StridedLaunch<T>>> m;
#define SH_V(n_, P_, C_)                          \
    m[n_].names.push_back(#P_ " columns=" #C_);   \
    m[n_].fns.push_back(&launch_strided<T, P_, C_>);
        SH_STRIDED_VARIANTS(SH_V)
#undef SH_V
#define SH_VT(n_, P_, C_)                                             \
    m[n_].names.push_back(#P_ " columns=" #C_ " nontemporal stores"); \
    m[n_].fns.push_back(&launch_strided<T, P_, C_, true>);
        SH_STRIDED_NT_VARIANTS(SH_VT)
#undef SH_VT
#define SH_VL(n_, P_, C_, NT_)                                                                    \
    m[n_].names.push_back(#P_ " columns=" #C_ " nontemporal loads, nontemporal stores=" #NT_);    \
    m[n_].fns.push_back(&launch_strided<T, P_, C_, NT_, true>);
        SH_STRIDED_LNT_VARIANTS(SH_VL)
#undef SH_VL
#define SH_VF(n_, P_, C_, NT_, LNT_)                                                                  \
    m[n_].names.push_back(#P_ " columns=" #C_ " nontemporal stores=" #NT_ ", nontemporal loads=" #LNT_); \
    m[n_].fns.push_back(&launch_strided<T, P_, C_, NT_, LNT_>);
        if constexpr (sizeof(T) == 4) {
            SH_STRIDED_F32_VARIANTS(SH_VF)
        }
#undef SH_VF
        return m;
    }();
    static const VariantList<StridedLaunch<T>> none;
    auto it = tab.find(n);
    return it == tab.end() ? none : it->second;
}

template <typename T>
constexpr int prec_of() {
    return sizeof(T) == 8 ? SH_F64 : SH_F32;
}

template <typename T>
int pass_contig(C2<T>* data, int n, long nlines, long ldist, int count, long fstride, int dir,
                const Epi<T>& ep, hipStream_t st, int role) {
    const auto& vl = contig_variants<T>(n);
    if (!vl.fns.empty() && role >= 0) {
        const int v = tune_choice(prec_of<T>(), n, role);
        if (v >= 0 && v < (int)vl.fns.size()) return vl.fns[v](data, nlines, ldist, count, fstride, dir, ep, st);
    }
    switch (n) {
#define SH_CASE(n_, P_, L_, C_) \
    case n_:                    \
        return launch_contig<T, P_, L_>(data, nlines, ldist, count, fstride, dir, ep, st);
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
        default:
            return pass_any<T>(data, n, nlines, nlines, ldist, 0, 1, count, fstride, dir, ep, false, st);
    }
}

template <typename T>
int pass_strided(C2<T>* data, int n, int ncols, long cdist, int nouter, long odist, long es,
                 int count, long fstride, int dir, const Epi<T>& ep, hipStream_t st, int role) {
    const auto& vl = strided_variants<T>(n);
    if (!vl.fns.empty() && role >= 0) {
        const int v = tune_choice(prec_of<T>(), n, role);
        if (v >= 0 && v < (int)vl.fns.size())
            return vl.fns[v](data, ncols, cdist, nouter, odist, es, count, fstride, dir, ep, st);
    }
    switch (n) {
#define SH_CASE(n_, P_, L_, C_)                                                                 \
    case n_:                                                                                    \
        return launch_strided<T, P_, C_>(data, ncols, cdist, nouter, odist, es, count, fstride, \
                                         dir, ep, st);
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
        default:
            return pass_any<T>(data, n, (long)ncols * nouter, ncols, cdist, odist, es, count, fstride, dir, ep, true,
                               st);
    }
}

template int pass_contig<double>(double2*, int, long, long, int, long, int, const Epi<double>&,
                                 hipStream_t, int);
template int pass_contig<float>(float2*, int, long, long, int, long, int, const Epi<float>&,
                                hipStream_t, int);
template int pass_strided<double>(double2*, int, int, long, int, long, long, int, long, int,
                                  const Epi<double>&, hipStream_t, int);
template int pass_strided<float>(float2*, int, int, long, int, long, long, int, long, int,
                                 const Epi<float>&, hipStream_t, int);

// number and names of the variants of a pass role at length n
int pass_variant_count(int precision, int n, int role) {
    if (role == ROLE_X)
        return precision == SH_F64 ? (int)contig_variants<double>(n).fns.size()
                                   : (int)contig_variants<float>(n).fns.size();
    return precision == SH_F64 ? (int)strided_variants<double>(n).fns.size()
                               : (int)strided_variants<float>(n).fns.size();
}

const char* pass_variant_name(int precision, int n, int role, int v) {
    if (role == ROLE_X) {
        const auto& names = precision == SH_F64 ? contig_variants<double>(n).names : contig_variants<float>(n).names;
        return v >= 0 && v < (int)names.size() ? names[v] : "default";
    }
    const auto& names = precision == SH_F64 ? strided_variants<double>(n).names : strided_variants<float>(n).names;
    return v >= 0 && v < (int)names.size() ? names[v] : "default";
}

template <typename T>
static int fft1d(void* data, int n, long stride, long dist, long batch, int dir, void* stream) {
    if (!data || n <= 0 || batch < 0 || stride <= 0 || (dir != 1 && dir != -1)) {
        set_error("sh_fft1d: invalid argument");
        return SH_EINVAL;
    }
    if (!fft_length_supported(n)) {
        set_error("sh_fft1d: unsupported length " + std::to_string(n));
        return SH_EUNSUPPORTED;
    }
    if (batch == 0) return SH_OK;
    int rc = ensure_init();
    if (rc) return rc;
    Epi<T> ep{T(1), 0, 1, -1, -1};
    hipStream_t st = as_stream(stream);
    if (stride == 1) return pass_contig<T>((C2<T>*)data, n, batch, dist, 1, 0, dir, ep, st, -1);
    if (batch > 0x7fffffffL) {
        set_error("sh_fft1d: batch too large for a strided transform");
        return SH_EINVAL;
    }
    return pass_strided<T>((C2<T>*)data, n, (int)batch, dist, 1, 0, stride, 1, 0, dir, ep, st, -1);
}

// 3D transform of `count` fields: x pass, y pass, z pass (+ epilogue).
template <typename T>
int fft3d(void* data, int count, long fstride, int nx, int ny, int nz, int dir, int flags,
          void* stream) {
    if (!data || count <= 0 || nx <= 0 || ny <= 0 || nz <= 0 || (dir != 1 && dir != -1)) {
        set_error("sh_fft3d: invalid argument");
        return SH_EINVAL;
    }
    if (!fft_length_supported(nx) || !fft_length_supported(ny) || !fft_length_supported(nz)) {
        set_error("sh_fft3d: unsupported dimensions " + std::to_string(nx) + "x" +
                  std::to_string(ny) + "x" + std::to_string(nz));
        return SH_EUNSUPPORTED;
    }
    int rc = ensure_init();
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    C2<T>* f = (C2<T>*)data;
    const long sxy = (long)nx * ny;
    Epi<T> plain{T(1), 0, 1, -1, -1};
    {
        ProfScope ps("fft_x", st);
        rc = pass_contig<T>(f, nx, (long)ny * nz, nx, count, fstride, dir, plain, st, ROLE_X);
    }
    if (rc) return rc;
    {
        ProfScope ps("fft_y", st);
        rc = pass_strided<T>(f, ny, nx, 1, nz, sxy, nx, count, fstride, dir, plain, st, ROLE_Y);
    }
    if (rc) return rc;
    Epi<T> last{T(1), 0, 1, nx / 2, ny / 2};
    if (flags & SH_NORMALIZE) {
        last.flags |= 1;
        last.scale = T(1.0 / ((double)nx * (double)ny * (double)nz));
    }
    if (flags & SH_ZERO_NYQUIST) last.flags |= 2;
    {
        ProfScope ps("fft_z", st);
        rc = pass_strided<T>(f, nz, nx, 1, ny, nx, sxy, count, fstride, dir, last, st, ROLE_Z);
    }
    return rc;
}

template int fft3d<double>(void*, int, long, int, int, int, int, int, void*);
template int fft3d<float>(void*, int, long, int, int, int, int, int, void*);

}  // namespace sh

using namespace sh;

extern "C" {

int sh_fft1d_z2z(void* data, int n, long stride, long dist, long batch, int dir, void* stream) {
    return fft1d<double>(data, n, stride, dist, batch, dir, stream);
}
int sh_fft1d_c2c(void* data, int n, long stride, long dist, long batch, int dir, void* stream) {
    return fft1d<float>(data, n, stride, dist, batch, dir, stream);
}
int sh_fft3d_z2z(void* data, int nx, int ny, int nz, int dir, int flags, void* stream) {
    return fft3d<double>(data, 1, 0, nx, ny, nz, dir, flags, stream);
}
int sh_fft3d_c2c(void* data, int nx, int ny, int nz, int dir, int flags, void* stream) {
    return fft3d<float>(data, 1, 0, nx, ny, nz, dir, flags, stream);
}
int sh_fft3d_many_z2z(void* data, int count, long field_stride, int nx, int ny, int nz, int dir,
                      int flags, void* stream) {
    return fft3d<double>(data, count, field_stride, nx, ny, nz, dir, flags, stream);
}
int sh_fft3d_many_c2c(void* data, int count, long field_stride, int nx, int ny, int nz, int dir,
                      int flags, void* stream) {
    return fft3d<float>(data, count, field_stride, nx, ny, nz, dir, flags, stream);
}

}  // extern "C"
