// Coulomb matrix elements on the device: pair densities B_ab(q) by FFT correlation of the
// envelope coefficients, upsampling to the Coulomb cell, the Poisson kernel and the
// contractions V = sum_q B_ij(q) w(q) B_kl(-q).
//
// Reference (code/source/8_band): coulomb_matrix_elements.c:24-627, state_redistribution.c,
// B_convolutions.c:332-430, B_coulomb_supercell.c:23-100, coulomb_integrals.c:23-356.
//
// Design.  The reference forms, per pair and band, FFT(conj-reversed c_a) * FFT(c_b) and one
// inverse FFT per pair and band on the (2N-1)^3 grid.  Here every state x band is transformed
// once on a power-of-two grid L >= 2N-1 per axis (zero padding keeps the linear correlation
// exact), the conj-reversed transform is replaced by the correlation theorem
// (FFT(conj-reversed c) = phase * conj(FFT(c)), the phase only shifts the output index), the
// band sum is taken in frequency space, and each pair needs a single inverse FFT.  The FFTs are
// radix-2 passes along one axis at a time, a tile of lines per work-group staged in LDS, with
// the tile chosen so that global accesses are contiguous runs.  The contraction of all pairs
// with all (reversed, kernel-weighted) pairs and the multipole vectors is one split-K complex
// FP64 product with 64x64 output tiles.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qd_coulomb.hpp"

namespace qd {

namespace {

constexpr int kFftLds = 2048;     // complex elements per work-group tile
constexpr int kFftThreads = 256;

__device__ inline c64 cmuls(c64 a, c64 b) { return a * b; }

// One radix-2 DIT pass of length L = 2^logL over all lines of one axis.
// Element m of line (g, outer, inner) sits at g*gstride + outer*L*IC + m*IC + inner.
template <bool kInverse>
__global__ __launch_bounds__(kFftThreads) void k_fft_axis(c64 *__restrict__ data, int logL, int IC, int OC,
                                                          long gstride, long nlines, const c64 *__restrict__ tw) {
    __shared__ c64 buf[kFftLds];
    __shared__ c64 stw[kFftLds / 2];
    const int L = 1 << logL;
    const int T = kFftLds >> logL;                  // lines in this tile
    const long first = (long)blockIdx.x * T;
    const int tid = threadIdx.x;
    for (int i = tid; i < L / 2; i += kFftThreads) {
        c64 w = tw[i];
        if (kInverse) w.im = -w.im;
        stw[i] = w;
    }
    auto addr = [&](long F, int m) -> long {
        const long inner = F % IC;
        const long rest = F / IC;
        const long outer = rest % OC, g = rest / OC;
        return g * gstride + outer * ((long)L * IC) + (long)m * IC + inner;
    };
    const int n = T * L;
    for (int e = tid; e < n; e += kFftThreads) {
        int j, m;
        if (IC == 1) { j = e >> logL; m = e & (L - 1); }
        else { m = e / T; j = e - m * T; }
        const long F = first + j;
        const int rm = (int)(__brev((unsigned)m) >> (32 - logL));
        buf[j * L + rm] = F < nlines ? data[addr(F, m)] : mk(0.0, 0.0);
    }
    __syncthreads();
    for (int s = 1; s <= logL; ++s) {
        const int half = 1 << (s - 1);
        const int step = L >> s;
        for (int b = tid; b < n / 2; b += kFftThreads) {
            const int j = b >> (logL - 1);
            const int r = b & (L / 2 - 1);
            const int pos = r & (half - 1);
            const int i0 = j * L + ((r >> (s - 1)) << s) + pos;
            const int i1 = i0 + half;
            const c64 u = buf[i0];
            const c64 v = cmuls(buf[i1], stw[pos * step]);
            buf[i0] = u + v;
            buf[i1] = u - v;
        }
        __syncthreads();
    }
    for (int e = tid; e < n; e += kFftThreads) {
        int j, m;
        if (IC == 1) { j = e >> logL; m = e & (L - 1); }
        else { m = e / T; j = e - m * T; }
        const long F = first + j;
        if (F < nlines) data[addr(F, m)] = buf[j * L + m];
    }
}

// Zero-padded grids of every state x band: F[(s*8+b)][x][y][z] = c_s[b*NG + (x*Ny+y)*Nz+z].
__global__ void k_pad(const c64 *__restrict__ X, long ldx, int S, int Nx, int Ny, int Nz, int Lx, int Ly, int Lz,
                      c64 *__restrict__ F) {
    const long grid = (long)Lx * Ly * Lz;
    const long total = grid * S * 8;
    const long NG = (long)Nx * Ny * Nz;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long gi = i / grid;
        const long r = i - gi * grid;
        const int x = (int)(r / ((long)Ly * Lz));
        const int y = (int)((r / Lz) % Ly);
        const int z = (int)(r % Lz);
        const int s = (int)(gi >> 3), b = (int)(gi & 7);
        c64 v = mk(0.0, 0.0);
        if (x < Nx && y < Ny && z < Nz) v = X[(long)s * ldx + b * NG + ((long)x * Ny + y) * Nz + z];
        F[i] = v;
    }
}

// Band-summed cross spectrum of each pair: Pg[p][f] = sum_b conj(F[a,b][f]) F[c,b][f].
__global__ void k_pair_spectrum(const c64 *__restrict__ F, long grid, const int2 *__restrict__ pairs,
                                c64 *__restrict__ Pg) {
    const int p = blockIdx.y;
    const int2 ac = pairs[p];
    const c64 *Fa = F + (long)ac.x * 8 * grid;
    const c64 *Fc = F + (long)ac.y * 8 * grid;
    for (long f = (long)blockIdx.x * blockDim.x + threadIdx.x; f < grid; f += (long)gridDim.x * blockDim.x) {
        c64 acc = mk(0.0, 0.0);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const c64 x = Fa[b * grid + f], y = Fc[b * grid + f];
            acc = acc + cmul(x, y);
        }
        Pg[(long)p * grid + f] = acc;
    }
}

// B[p][n] on the (2N-1)^3 grid: the cyclic correlation at q = n - (N-1), scaled by 1/|grid|.
__global__ void k_extract(const c64 *__restrict__ Pg, int P, int Lx, int Ly, int Lz, int Nx, int Ny, int Nz,
                          double scale, c64 *__restrict__ B) {
    const int NTx = 2 * Nx - 1, NTy = 2 * Ny - 1, NTz = 2 * Nz - 1;
    const long nes = (long)NTx * NTy * NTz;
    const long grid = (long)Lx * Ly * Lz;
    const long total = nes * P;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long p = i / nes;
        const long n = i - p * nes;
        const int ix = (int)(n / ((long)NTy * NTz)), iy = (int)((n / NTz) % NTy), iz = (int)(n % NTz);
        int qx = ix - (Nx - 1), qy = iy - (Ny - 1), qz = iz - (Nz - 1);
        qx += qx < 0 ? Lx : 0;
        qy += qy < 0 ? Ly : 0;
        qz += qz < 0 ? Lz : 0;
        const c64 v = Pg[p * grid + ((long)qx * Ly + qy) * Lz + qz];
        B[i] = scale * v;
    }
}

// out[..., i, ...] = sum_j M[i][j] in[..., j, ...] along one axis of a batch of 3-D grids.
__global__ void k_apply_axis(const c64 *__restrict__ in, int d0, int d1, int d2, int axis, int m,
                             const double *__restrict__ M, long batch, c64 *__restrict__ out) {
    int o0 = d0, o1 = d1, o2 = d2;
    if (axis == 0) o0 = m;
    else if (axis == 1) o1 = m;
    else o2 = m;
    const long osz = (long)o0 * o1 * o2, isz = (long)d0 * d1 * d2;
    const long total = osz * batch;
    const int nin = axis == 0 ? d0 : (axis == 1 ? d1 : d2);
    const long istride = axis == 0 ? (long)d1 * d2 : (axis == 1 ? d2 : 1);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long b = t / osz;
        const long r = t - b * osz;
        const int x = (int)(r / ((long)o1 * o2)), y = (int)((r / o2) % o1), z = (int)(r % o2);
        const int i = axis == 0 ? x : (axis == 1 ? y : z);
        // input index of (x, y, z) with the transformed coordinate set to 0
        long b0 = b * isz;
        if (axis == 0) b0 += (long)y * d2 + z;
        else if (axis == 1) b0 += (long)x * d1 * d2 + z;
        else b0 += ((long)x * d1 + y) * d2;
        c64 acc = mk(0.0, 0.0);
        const double *Mi = M + (long)i * nin;
        for (int j = 0; j < nin; ++j) acc = acc + Mi[j] * in[b0 + j * istride];
        out[t] = acc;
    }
}

// Contraction operand rows: Y[r] = B_r(-q) w(q) (r < P), conj(B_{r-P}(-q)) w(q) (P <= r < 2P),
// then the uploaded multipole rows.
__global__ void k_make_y(const c64 *__restrict__ B, int P, long nq, const double *__restrict__ w,
                         const c64 *__restrict__ vec, int nvec, c64 *__restrict__ Y) {
    const long total = (long)(2 * P + nvec) * nq;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long r = t / nq;
        const long n = t - r * nq;
        c64 v;
        if (r < 2 * P) {
            const long p = r < P ? r : r - P;
            v = w[n] * B[p * nq + (nq - 1 - n)];
            if (r >= P) v.im = -v.im;
        } else {
            v = vec[(r - 2 * P) * nq + n];
        }
        Y[t] = v;
    }
}

constexpr int kCT = 64;    // output tile (pairs x rows)
constexpr int kCK = 32;    // K slice staged in LDS

// Split-K partial products part[kc][p][r] = sum_{n in chunk kc} A[p][n] Y[r][n].
__global__ __launch_bounds__(256) void k_contract(const c64 *__restrict__ A, int P, const c64 *__restrict__ Y, int R,
                                                  long nq, long chunk, c64 *__restrict__ part) {
    __shared__ c64 sA[kCT][kCK + 1];
    __shared__ c64 sY[kCT][kCK + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int p0 = blockIdx.x * kCT, r0 = blockIdx.y * kCT;
    const long k0 = (long)blockIdx.z * chunk;
    const long k1 = std::min(nq, k0 + chunk);
    c64 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mk(0.0, 0.0);
    for (long kb = k0; kb < k1; kb += kCK) {
        for (int e = threadIdx.x; e < kCT * kCK; e += 256) {
            const int row = e / kCK, kk = e % kCK;
            const long k = kb + kk;
            const int p = p0 + row, r = r0 + row;
            sA[row][kk] = (p < P && k < k1) ? A[(long)p * nq + k] : mk(0.0, 0.0);
            sY[row][kk] = (r < R && k < k1) ? Y[(long)r * nq + k] : mk(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < kCK; ++kk) {
            c64 a[4], y[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sA[ty + 16 * i][kk];
#pragma unroll
            for (int j = 0; j < 4; ++j) y[j] = sY[tx + 16 * j][kk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) fma_(acc[i][j], a[i], y[j]);
        }
        __syncthreads();
    }
    c64 *out = part + (long)blockIdx.z * P * R;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int p = p0 + ty + 16 * i, r = r0 + tx + 16 * j;
            if (p < P && r < R) out[(long)p * R + r] = acc[i][j];
        }
}

// C[t] = sum over the split-K chunks of part[c][t]: 16 outputs x 16 chunk groups per work-group, four
// loads in flight per thread (clamped, masked), the groups added in LDS
__global__ __launch_bounds__(256) void k_sum_chunks(const c64 *__restrict__ part, int nchunk, long n, c64 *__restrict__ C) {
    __shared__ c64 red[16][17];
    const int o = threadIdx.x & 15, g = threadIdx.x >> 4;
    const long t = (long)blockIdx.x * 16 + o;
    const long tc = t < n ? t : n - 1;
    c64 acc = mk(0.0, 0.0);
    for (int c0 = g; c0 < nchunk; c0 += 64) {
        c64 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = part[(long)min(c0 + 16 * u, nchunk - 1) * n + tc];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (c0 + 16 * u < nchunk) acc = acc + v[u];
    }
    red[g][o] = acc;
    __syncthreads();
    if (g == 0 && t < n) {
        c64 sum = mk(0.0, 0.0);
        for (int q = 0; q < 16; ++q) sum = sum + red[q][o];
        C[t] = sum;
    }
}

unsigned blocks_for(long n, int threads = 256, long cap = 1 << 16) {
    return (unsigned)std::max<long>(1, std::min<long>(cap, (n + threads - 1) / threads));
}

}  // namespace

int ilog2_ceil(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

void apply_axis_device(const c64 *in, const int dims[3], int axis, int m, const double *M, long batch, c64 *out,
                       hipStream_t s) {
    long osz = batch * m;
    for (int a = 0; a < 3; ++a)
        if (a != axis) osz *= dims[a];
    hipLaunchKernelGGL(k_apply_axis, dim3(blocks_for(osz)), dim3(256), 0, s, in, dims[0], dims[1], dims[2], axis, m, M,
                       batch, out);
    QD_LAUNCH_CHECK();
}

void fft3_device(c64 *data, long ngrids, const int logL[3], const c64 *const tw[3], bool inverse, hipStream_t s) {
    const int L[3] = {1 << logL[0], 1 << logL[1], 1 << logL[2]};
    const long grid = (long)L[0] * L[1] * L[2];
    for (int a = 2; a >= 0; --a) {
        const int IC = a == 2 ? 1 : (a == 1 ? L[2] : L[1] * L[2]);
        const int OC = (int)(grid / ((long)L[a] * IC));
        const long nlines = ngrids * OC * IC;
        const int T = kFftLds >> logL[a];
        const unsigned nb = (unsigned)((nlines + T - 1) / T);
        if (inverse)
            hipLaunchKernelGGL(k_fft_axis<true>, dim3(nb), dim3(kFftThreads), 0, s, data, logL[a], IC, OC, grid, nlines, tw[a]);
        else
            hipLaunchKernelGGL(k_fft_axis<false>, dim3(nb), dim3(kFftThreads), 0, s, data, logL[a], IC, OC, grid, nlines, tw[a]);
        QD_LAUNCH_CHECK();
    }
}


namespace {

constexpr double kESQEPS = 180.9513;   // e^2/eps0 [eV Angstrom], 8_band.h:52

__device__ inline double block_sum_d(double v, double *sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

// I_a(g) = CLb CLc int_{-CLa/2}^{CLa/2} x e^{i g x} dx = i CLb CLc (2 sin(g CLa/2) / g^2 - CLa cos(g CLa/2) / g)
// (coulomb_matrix_elements.c:408-424)
__device__ inline c64 dipole_axis(double g, double CLa, double CLb, double CLc) {
    const double h = 0.5 * g * CLa;
    return mk(0.0, CLb * CLc * (2.0 * sin(h) / (g * g) - CLa * cos(h) / g));
}

// J_a(g) = CLb CLc int x^2 e^{i g x} dx = CLb CLc (CLa^2 sin(h) / (2 g) + 2 CLa cos(h) / g^2 - 4 sin(h) / g^3),
// h = g CLa / 2 (:435-478)
__device__ inline double quad_axis(double g, double CLa, double CLb, double CLc) {
    const double h = 0.5 * g * CLa;
    return CLb * CLc * (CLa * CLa * sin(h) / (2.0 * g) + 2.0 * CLa * cos(h) / (g * g) - 4.0 * sin(h) / (g * g * g));
}

// The Coulomb-cell vectors at every q of the Coulomb grid (coulomb_matrix_elements.c:384-478): the
// Poisson kernel w(q) = ESQEPS / (Omega_C |q|^2 nr0^2) (0 at q = 0); for CME >= 3 the dipole vectors
// I_x, I_y, I_z (nonzero on their axis through the centre), their conjugates and the quadrupole
// vector J (CME >= 4, with its q = 0 value), rows 0..6 of vec; and per work-group sums of the
// reciprocal-space Ewald term (4 pi / Omega_C) exp(-|q|^2 / (4 eta^2)) / |q|^2 of the Madelung
// constant (:250-334).
__global__ __launch_bounds__(256) void k_cme_vectors(const double *__restrict__ gx, const double *__restrict__ gy,
                                                     const double *__restrict__ gz, int Nx, int Ny, int Nz, double CLx,
                                                     double CLy, double CLz, double omega_c, double nr0, double eta,
                                                     int cme, long nq, double *__restrict__ w, c64 *__restrict__ vec,
                                                     double *__restrict__ rec_part) {
    __shared__ double sh[256];
    const double four_pi = 4.0 * M_PI;
    double rec = 0.0;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < nq; q += (long)gridDim.x * 256) {
        const int ix = (int)(q / ((long)Ny * Nz)), iy = (int)((q / Nz) % Ny), iz = (int)(q % Nz);
        const double G2 = gx[ix] * gx[ix] + gy[iy] * gy[iy] + gz[iz] * gz[iz];
        const bool cx = ix == Nx / 2, cy = iy == Ny / 2, cz = iz == Nz / 2, cen = cx && cy && cz;
        w[q] = cen ? 0.0 : kESQEPS / (omega_c * G2 * nr0 * nr0);
        if (!cen) rec += (four_pi / omega_c) * exp(-G2 / (4.0 * eta * eta)) / G2;
        if (cme >= 3) {
            c64 I[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
            double J = 0.0;
            if (!cx && cy && cz) I[0] = dipole_axis(gx[ix], CLx, CLy, CLz);
            if (cx && !cy && cz) I[1] = dipole_axis(gy[iy], CLy, CLx, CLz);
            if (cx && cy && !cz) I[2] = dipole_axis(gz[iz], CLz, CLy, CLx);
            if (cme >= 4) {
                if (!cx && cy && cz) J += quad_axis(gx[ix], CLx, CLy, CLz);
                if (cx && !cy && cz) J += quad_axis(gy[iy], CLy, CLx, CLz);
                if (cx && cy && !cz) J += quad_axis(gz[iz], CLz, CLy, CLx);
                if (cen) J += (CLx * CLx * CLx * CLy * CLz + CLy * CLy * CLy * CLx * CLz + CLz * CLz * CLz * CLy * CLx) / 12.0;
            }
            for (int a = 0; a < 3; ++a) {
                vec[a * nq + q] = I[a];
                vec[(3 + a) * nq + q] = conj(I[a]);
            }
            vec[6 * nq + q] = mk(J, 0.0);
        }
    }
    const double t = block_sum_d(rec, sh);
    if (threadIdx.x == 0) rec_part[blockIdx.x] = t;
}

// Per work-group sums of the real-space Ewald term erfc(eta R) / R over the flat lattice indices
// idx < hi, (mx, my, mz) = (idx / (hy hz), (idx mod hy hz) / hz, idx mod hz), R = |(mx CLx, my CLy,
// mz CLz)|, the origin excluded (coulomb_matrix_elements.c:287-305).
__global__ __launch_bounds__(256) void k_madelung_real(long hi, long hy, long hz, double CLx, double CLy, double CLz,
                                                       double eta, double *__restrict__ part) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < hi; idx += (long)gridDim.x * 256) {
        const long mx = idx / (hy * hz), my = (idx % (hy * hz)) / hz, mz = (idx % (hy * hz)) % hz;
        if (mx == 0 && my == 0 && mz == 0) continue;
        const double X = (double)mx * CLx, Y = (double)my * CLy, Z = (double)mz * CLz;
        const double R = sqrt(X * X + Y * Y + Z * Z);
        acc += erfc(R * eta) / R;
    }
    const double t = block_sum_d(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Device buffers of one Coulomb evaluation, carved from one arena (256-byte aligned pieces).
// With base == nullptr the same walk only sizes the arena.
struct Carver {
    unsigned char *base = nullptr;
    size_t cap = 0, used = 0;
    template <class T>
    T *take(size_t n) {
        const size_t off = (used + 255) & ~size_t(255);
        const size_t bytes = sizeof(T) * std::max<size_t>(n, 1);
        if (base && off + bytes > cap) throw StatusError("Coulomb scratch smaller than its layout", QD_E_ARG);
        used = off + bytes;
        return base ? reinterpret_cast<T *>(base + off) : nullptr;
    }
};

constexpr int kEwaldBlocks = 1024;   // work-groups (and partial sums) of each Ewald sum

struct CoulombBufs {
    c64 *tw[3], *F, *Pg, *B, *U[3], *Y, *vd, *part, *Cd;
    double *Md[3], *wd, *gq[3], *ewald;   // ewald: 2 x kEwaldBlocks partial sums
    int2 *pairs;
    int P, R, nvec, logL[3], fft[3];
    long grid, nchunk, chunk;
};

void coulomb_layout(int S, const CoulombSetup &cs, Carver &a, CoulombBufs &b) {
    b.P = S * (S + 1) / 2;
    b.nvec = cs.cme >= 3 ? 7 : 0;
    b.R = 2 * b.P + b.nvec;
    for (int k = 0; k < 3; ++k) {
        b.logL[k] = std::max(1, ilog2_ceil(cs.NT[k]));
        if (b.logL[k] > 9) throw HipError("Coulomb FFT grid above 512 points per axis", QD_E_UNSUPPORTED);
        b.fft[k] = 1 << b.logL[k];
    }
    b.grid = (long)b.fft[0] * b.fft[1] * b.fft[2];
    const long nes = cs.nes(), nq = cs.nq();
    for (int k = 0; k < 3; ++k) b.tw[k] = a.take<c64>(std::max(1, b.fft[k] / 2));
    b.pairs = a.take<int2>(b.P);
    b.F = a.take<c64>((size_t)S * 8 * b.grid);
    b.Pg = a.take<c64>((size_t)b.P * b.grid);
    b.B = a.take<c64>((size_t)b.P * nes);
    int dims[3] = {cs.NT[0], cs.NT[1], cs.NT[2]};
    for (int k = 0; k < 3; ++k) {
        b.U[k] = nullptr;
        b.Md[k] = nullptr;
        if (cs.identity() || cs.copy[k]) continue;
        b.Md[k] = a.take<double>(cs.up[k].size());
        dims[k] = cs.NTC[k];
        b.U[k] = a.take<c64>((size_t)dims[0] * dims[1] * dims[2] * b.P);
    }
    b.wd = a.take<double>(nq);
    for (int k = 0; k < 3; ++k) b.gq[k] = a.take<double>(cs.NTC[k]);
    b.ewald = a.take<double>(2 * kEwaldBlocks);
    b.vd = a.take<c64>(b.nvec ? (size_t)b.nvec * nq : 1);
    b.Y = a.take<c64>((size_t)b.R * nq);
    const int tp = (b.P + kCT - 1) / kCT, tr = (b.R + kCT - 1) / kCT;
    long nchunk = std::max<long>(1, std::min<long>((nq + kCK - 1) / kCK, 1024 / std::max(1, tp * tr)));
    long chunk = (nq + nchunk - 1) / nchunk;
    chunk = (chunk + kCK - 1) / kCK * kCK;
    b.nchunk = (nq + chunk - 1) / chunk;
    b.chunk = chunk;
    b.part = a.take<c64>((size_t)b.nchunk * b.P * b.R);
    b.Cd = a.take<c64>((size_t)b.P * b.R);
}

}  // namespace

size_t coulomb_scratch_bytes(int S, const CoulombSetup &cs) {
    Carver a;
    CoulombBufs b;
    coulomb_layout(S, cs, a, b);
    return a.used;
}

void coulomb_device(const c64 *X, long ldx, int S, const CoulombSetup &cs, unsigned char *scratch,
                    size_t scratch_bytes, const hipEvent_t ev[3], hipStream_t s, std::vector<zc> &Cout, int &P_out,
                    int &R_out, int fft_len[3], double *t_pairs, double *t_contract, double *a_mad) {
    Carver a;
    a.base = scratch;
    a.cap = scratch_bytes;
    CoulombBufs b;
    coulomb_layout(S, cs, a, b);
    const int P = b.P, R = b.R, nvec = b.nvec;
    const long grid = b.grid, nes = cs.nes(), nq = cs.nq();
    for (int k = 0; k < 3; ++k) fft_len[k] = b.fft[k];
    // twiddles exp(-2 pi i k / L), k < L/2, per axis
    const c64 *tw[3];
    for (int k = 0; k < 3; ++k) {
        const int L = b.fft[k];
        std::vector<c64> h(std::max(1, L / 2));
        for (int j = 0; j < L / 2; ++j) {
            const double ang = -2.0 * M_PI * (double)j / (double)L;
            h[j] = mk(std::cos(ang), std::sin(ang));
        }
        QD_HIP(hipMemcpyAsync(b.tw[k], h.data(), sizeof(c64) * h.size(), hipMemcpyHostToDevice, s));
        tw[k] = b.tw[k];
    }
    std::vector<int2> hpairs(P);
    for (int j = 0; j < S; ++j)
        for (int i = 0; i <= j; ++i) hpairs[pair_index(i, j)] = make_int2(i, j);
    QD_HIP(hipMemcpyAsync(b.pairs, hpairs.data(), sizeof(int2) * P, hipMemcpyHostToDevice, s));

    QD_HIP(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL(k_pad, dim3(blocks_for((long)S * 8 * grid)), dim3(256), 0, s, X, ldx, S, cs.N[0], cs.N[1],
                       cs.N[2], fft_len[0], fft_len[1], fft_len[2], b.F);
    QD_LAUNCH_CHECK();
    fft3_device(b.F, (long)S * 8, b.logL, tw, false, s);
    hipLaunchKernelGGL(k_pair_spectrum, dim3(blocks_for(grid, 256, 4096), P), dim3(256), 0, s, b.F, grid, b.pairs, b.Pg);
    QD_LAUNCH_CHECK();
    fft3_device(b.Pg, P, b.logL, tw, true, s);
    hipLaunchKernelGGL(k_extract, dim3(blocks_for((long)P * nes)), dim3(256), 0, s, b.Pg, P, fft_len[0], fft_len[1],
                       fft_len[2], cs.N[0], cs.N[1], cs.N[2], 1.0 / (double)grid, b.B);
    QD_LAUNCH_CHECK();
    QD_HIP(hipEventRecord(ev[1], s));

    // upsampling to the Coulomb cell (separable box transform), then contraction operands
    const c64 *BC = b.B;
    if (!cs.identity()) {
        int dims[3] = {cs.NT[0], cs.NT[1], cs.NT[2]};
        const c64 *cur = b.B;
        for (int k = 0; k < 3; ++k) {
            if (cs.copy[k]) continue;
            QD_HIP(hipMemcpyAsync(b.Md[k], cs.up[k].data(), sizeof(double) * cs.up[k].size(), hipMemcpyHostToDevice, s));
            int od[3] = {dims[0], dims[1], dims[2]};
            od[k] = cs.NTC[k];
            const long osz = (long)od[0] * od[1] * od[2] * P;
            hipLaunchKernelGGL(k_apply_axis, dim3(blocks_for(osz)), dim3(256), 0, s, cur, dims[0], dims[1], dims[2], k,
                               cs.NTC[k], b.Md[k], (long)P, b.U[k]);
            QD_LAUNCH_CHECK();
            cur = b.U[k];
            dims[k] = cs.NTC[k];
        }
        BC = cur;
    }
    // Poisson kernel, multipole vectors and the Ewald sums of the Madelung constant
    for (int k = 0; k < 3; ++k)
        QD_HIP(hipMemcpyAsync(b.gq[k], cs.gq[k].data(), sizeof(double) * cs.NTC[k], hipMemcpyHostToDevice, s));
    const unsigned nb_rec = (unsigned)std::min<long>((nq + 255) / 256, kEwaldBlocks);
    hipLaunchKernelGGL(k_cme_vectors, dim3(nb_rec), dim3(256), 0, s, b.gq[0], b.gq[1], b.gq[2], cs.NTC[0], cs.NTC[1],
                       cs.NTC[2], cs.CL[0], cs.CL[1], cs.CL[2], cs.omega_c, cs.nr0, cs.eta, cs.cme, nq, b.wd, b.vd,
                       b.ewald);
    QD_LAUNCH_CHECK();
    const unsigned nb_real = (unsigned)std::max<long>(1, std::min<long>((cs.mad_hi + 255) / 256, kEwaldBlocks));
    if (cs.cme >= 2) {
        hipLaunchKernelGGL(k_madelung_real, dim3(nb_real), dim3(256), 0, s, cs.mad_hi, std::max(1L, cs.mad_hy),
                           std::max(1L, cs.mad_hz), cs.CL[0], cs.CL[1], cs.CL[2], cs.eta, b.ewald + kEwaldBlocks);
        QD_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_make_y, dim3(blocks_for((long)R * nq)), dim3(256), 0, s, BC, P, nq, b.wd, b.vd, nvec, b.Y);
    QD_LAUNCH_CHECK();
    const int tp = (P + kCT - 1) / kCT, tr = (R + kCT - 1) / kCT;
    hipLaunchKernelGGL(k_contract, dim3(tp, tr, (unsigned)b.nchunk), dim3(256), 0, s, BC, P, b.Y, R, nq, b.chunk, b.part);
    QD_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_chunks, dim3((unsigned)(((long)P * R + 15) / 16)), dim3(256), 0, s, b.part, (int)b.nchunk,
                       (long)P * R, b.Cd);
    QD_LAUNCH_CHECK();
    QD_HIP(hipEventRecord(ev[2], s));
    Cout.resize((size_t)P * R);
    QD_HIP(hipMemcpyAsync(Cout.data(), b.Cd, sizeof(c64) * P * R, hipMemcpyDeviceToHost, s));
    std::vector<double> ew(2 * kEwaldBlocks, 0.0);
    QD_HIP(hipMemcpyAsync(ew.data(), b.ewald, sizeof(double) * nb_rec, hipMemcpyDeviceToHost, s));
    if (cs.cme >= 2)
        QD_HIP(hipMemcpyAsync(ew.data() + kEwaldBlocks, b.ewald + kEwaldBlocks, sizeof(double) * nb_real,
                              hipMemcpyDeviceToHost, s));
    QD_HIP(hipStreamSynchronize(s));
    *a_mad = 0.0;
    if (cs.cme >= 2) {
        double rec = 0.0, real = 0.0;
        for (unsigned i = 0; i < nb_rec; ++i) rec += ew[i];
        for (unsigned i = 0; i < nb_real; ++i) real += ew[kEwaldBlocks + i];
        *a_mad = madelung_constant(cs, rec, cs.mad_hi > 0 ? real : 0.0);
    }
    float ms1 = 0, ms2 = 0;
    QD_HIP(hipEventElapsedTime(&ms1, ev[0], ev[1]));
    QD_HIP(hipEventElapsedTime(&ms2, ev[1], ev[2]));
    *t_pairs = ms1 * 1e-3;
    *t_contract = ms2 * 1e-3;
    P_out = P;
    R_out = R;
}

}  // namespace qd
