// Nodal kernels and the consistent-mass solves.
//
// The mass matrix of a cartesian Q_p mesh is Mx (x) My, and the TG3 matrix
// M + c K_e adds Kronecker terms of the same 1-D bands, so a matrix-vector
// product is the tile pass of the advection kernel with other coefficients.
// CG runs for up to eight right-hand sides side by side (one per moving
// population, as the reference's eight sequential SolverCG solves): each
// vector keeps its own alpha, beta, residual and stopping test in device
// memory, and a converged vector is skipped from then on, so that every vector
// takes exactly the iterations its own solve would.
//
// Reductions: wavefront shuffles and LDS inside a block, one partial per block,
// then the partials summed in block order (a fixed order, so results do not
// depend on scheduling) and the scalars updated: by a one-block-per-vector
// finalisation launch after the start and direction stages, by the last
// update block of each vector to finish (counted with an atomic) after the
// update stage.
#include <algorithm>
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"

namespace lbfem_amd
{
  namespace
  {
    constexpr int kThreads = 256;
    constexpr int kPart    = kCgPartials; // partial sums per block, contiguous

    template <int p>
    __device__ __forceinline__ double
    band(const double *B, const int i, const int o)
    {
      return B[i * (2 * p + 1) + o + p];
    }

    __device__ __forceinline__ int
    wrap(const int i, const int P, const int periodic)
    {
      if (periodic)
        return ((i % P) + P) % P;
      return (i < 0 || i > P) ? -1 : i;
    }

    __device__ __forceinline__ bool
    is_slave(const Geo &G, const int i, const int j)
    {
      return (G.ax.periodic && i == G.P) || (G.ay.periodic && j == G.P);
    }

    // the Jacobi preconditioner: the inverse of diag(M) = diag(Mx) diag(My) per node, 1 on
    // constrained rows, applied as a product (DiagonalMatrix::vmult of the inverse entries)
    __device__ __forceinline__ double
    mass_diag_inv(const Geo &G, const int i, const int j)
    {
      return G.dinv[node_index(G, i, j)];
    }

    // the inverse lumped mass, 1 on constrained rows
    __device__ __forceinline__ double
    lumped_inv(const Geo &G, const int i, const int j)
    {
      if (G.gen.ml_inv)
        return G.gen.ml_inv[node_index(G, i, j)];
      return is_slave(G, i, j) ? 1. : G.ax.ml_inv[i] * G.ay.ml_inv[j];
    }

    template <int K>
    __device__ __forceinline__ void
    block_sum(double (&v)[K], double (*lds)[kThreads / 64])
    {
      block_sum_n<K, kThreads>(v, lds);
    }
  } // namespace

  // ------------------------------------------------------------------ nodal kernels
  __global__ void __launch_bounds__(kThreads)
  k_collide(const Geo G, double *g, const double omega)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        double    v[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = g[a * G.stride + k];
        double     wx = 0., wy = 0.;
        const bool wall = wall_of(G, i, j, wx, wy);
        collide_node(v, omega, wall, wx, wy);
#pragma unroll
        for (int a = 0; a < Q; ++a)
          g[a * G.stride + k] = v[a];
      }
  }

  __global__ void __launch_bounds__(kThreads)
  k_equilibrium(const Geo G, const double *f, double *feq)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        double    v[Q], e[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = f[a * G.stride + k];
        double     wx = 0., wy = 0.;
        const bool wall = wall_of(G, i, j, wx, wy);
        equilibrium_node(v, e, wall, wx, wy);
#pragma unroll
        for (int a = 0; a < Q; ++a)
          feq[a * G.stride + k] = e[a];
      }
  }

  __global__ void __launch_bounds__(kThreads)
  k_predictor_corrector(const Geo G, double *f, const double *feq, const double *incr, const double theta)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        double    v[Q], e[Q], x[NM];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          {
            v[a] = f[a * G.stride + k];
            e[a] = feq[a * G.stride + k];
          }
#pragma unroll
        for (int a = 0; a < NM; ++a)
          x[a] = incr[a * G.stride + k];
        double     wx = 0., wy = 0.;
        const bool wall = wall_of(G, i, j, wx, wy);
        predictor_corrector_node(v, e, x, theta, wall, wx, wy);
#pragma unroll
        for (int a = 0; a < Q; ++a)
          f[a * G.stride + k] = v[a];
      }
  }

  // f <- predictor-corrector(f, feq(f), x) in place: the Lee-Lin nodal update
  // after a consistent-mass solve, with the equilibria of f^n rebuilt
  // (collision.h:105-145)
  __global__ void __launch_bounds__(kThreads)
  k_pc_from_f(const Geo G, double *f, const double *incr, const double theta)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        double    v[Q], e[Q], x[NM];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = f[a * G.stride + k];
#pragma unroll
        for (int a = 0; a < NM; ++a)
          x[a] = incr[a * G.stride + k];
        double     wx = 0., wy = 0.;
        const bool wall = wall_of(G, i, j, wx, wy);
        equilibrium_node(v, e, wall, wx, wy);
        predictor_corrector_node(v, e, x, theta, wall, wx, wy);
#pragma unroll
        for (int a = 0; a < Q; ++a)
          f[a * G.stride + k] = v[a];
      }
  }

  __global__ void __launch_bounds__(kThreads)
  k_add_increments(const Geo G, double *f, const double *x, const unsigned mask)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
#pragma unroll
      for (int a = 0; a < NM; ++a)
        if (mask & (1u << a))
          f[(a + 1) * G.stride + k] += x[a * G.stride + k];
  }

  template <typename V> // the FP64 form's increments, or the mixed form's FP32 ones
  __global__ void __launch_bounds__(kThreads)
  k_extrapolate(const Geo G, V *x, V *x_prev, const unsigned mask)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
#pragma unroll
      for (int a = 0; a < NM; ++a)
        if (mask & (1u << a))
          {
            const V xa = x[a * G.stride + k], xp = x_prev[a * G.stride + k];
            x[a * G.stride + k]      = V(-double(xp) + 2. * double(xa)); // sadd(-1, 2, xa), then swap
            x_prev[a * G.stride + k] = xa;
          }
  }

  __global__ void __launch_bounds__(kThreads)
  k_lumped_solve(const Geo G, double *x, const double *r, const int nv)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int    i = int(k % G.np), j = int(k / G.np);
        const double m = lumped_inv(G, i, j);
        for (int v = 0; v < nv; ++v)
          x[v * G.stride + k] = m * r[v * G.stride + k];
      }
  }

  __global__ void __launch_bounds__(kThreads)
  k_richardson_update(const Geo G, double *x, const double *r, const double *Ax, const int nv)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int    i = int(k % G.np), j = int(k / G.np);
        const double m = lumped_inv(G, i, j);
        for (int v = 0; v < nv; ++v)
          x[v * G.stride + k] += m * (r[v * G.stride + k] - Ax[v * G.stride + k]);
      }
  }

  // the stopping states after iteration k (0: the start)
  __device__ __forceinline__ CgState *
  states_of(const CgWork &W, const int k)
  {
    return (W.fold_rows > 0 && (k & 1)) ? W.state2 : W.state;
  }

  // A stopping state read across a wavefront's lanes, lane l its l-th 8-byte word: a vector
  // load, issued with the vector loads around it (a scalar load of it is issued late, at its
  // first use), the fields taken from the lanes once it has arrived
  constexpr int kStateWords = int(sizeof(CgState) / sizeof(double));
  static_assert(sizeof(CgState) == kStateWords * sizeof(double), "CgState in 8-byte words");

  __device__ __forceinline__ double
  state_word(const CgState *s)
  {
    return reinterpret_cast<const double *>(s)[std::min(int(threadIdx.x & 63), kStateWords - 1)];
  }

  __device__ __forceinline__ CgState
  state_from_lanes(const double w)
  {
    const uint64_t bits = __builtin_bit_cast(uint64_t, w);
    const int      lo = int(uint32_t(bits)), hi = int(uint32_t(bits >> 32));
    uint64_t       words[kStateWords];
#pragma unroll
    for (int f = 0; f < kStateWords; ++f)
      words[f] = uint64_t(uint32_t(__builtin_amdgcn_readlane(lo, f))) |
                 (uint64_t(uint32_t(__builtin_amdgcn_readlane(hi, f))) << 32);
    return __builtin_bit_cast(CgState, words);
  }

  // SolverCG's stopping test (SolverControl::check: success first, then failure at the
  // cap) and beta after iteration it, from |r|^2 and r.D^-1 r; a stopped vector keeps its state
  __device__ __forceinline__ CgState
  advance_state(CgState st, const double rr, const double rz, const int it, const int max_its)
  {
    if (st.status != 0)
      return st;
    st.res = sqrt(fabs(rr));
    st.its = it;
    if (st.res <= st.tol)
      st.status = 1;
    else if (it >= max_its || isnan(st.res))
      st.status = 2;
    st.rho_prev = st.rho;
    st.rho      = rz;
    st.beta     = st.rho / st.rho_prev;
    return st;
  }

  // ------------------------------------------------------------------ operator tile
  // KIND 0: y = A x.  KIND 1 (CG start): r = A x - b, partials |r|^2, |b|^2, r.D^-1 r.
  // KIND 2 (CG direction): p_new = beta p_old - D^-1 r, q = A p_new, partial p_new.q.
  // OP kOpMass: A = Mx (x) My (TG2 mass solves, Richardson, the vectors without a K_e term).
  // OP kOpStreaming: A = M + c K_e = Mx (My + c ey^2 Ky) + Kx (c ex^2 My) + Dx^T (c ex ey Dy)
  //   + Dx (c ex ey Dy^T) (mass.h:105-144; a vector with c = 0 or e = 0 takes the mass branch).
  // OP kOpLaplace: the stream function's A = Kx (x) My + Mx (x) Ky on a walled box (io.h:55-85):
  //   wall entries read as zero, wall rows the identity, the Jacobi diagonal from Geo::dinv.
  // A thread keeps only the 1-D bands its OP uses (registers set the occupancy).
  // One vector per block (blockIdx.y): its operand tile is staged in LDS (one barrier), then
  // each thread applies the bands at its RO owned nodes of one column.
  // V: the vector type, double (the FP64 solves) or float (the mixed form's solves);
  // scalars and partial sums in FP64.
  enum
  {
    kOpMass      = 0,
    kOpStreaming = 1,
    kOpLaplace   = 2
  };

  template <int OP>
  struct OpBands
  {
    // bands a thread holds: M; M, K; or M, D, K, D^T (slot of each, -1 when absent)
    static constexpr int NB = OP == kOpMass ? 1 : OP == kOpLaplace ? 2 : 4;
    static constexpr int M = 0, K = OP == kOpMass ? -1 : OP == kOpLaplace ? 1 : 2;
    static constexpr int D = OP == kOpStreaming ? 1 : -1, DT = OP == kOpStreaming ? 3 : -1;
  };

  __device__ __forceinline__ const double *
  band_array(const AxisDev &a, const int which) // 0 M, 1 D, 2 K, 3 D^T
  {
    return which == 0 ? a.M : which == 1 ? a.D : which == 2 ? a.K : a.DT;
  }

  // The operator's tile: HX = 64 halo'd columns, one wavefront's lanes (lane l stages and
  // reads column l of every row), TX = 64 - 2p owned columns, TY = 16 owned rows; wavefront
  // w stages halo rows w, w + 4, ... and owns rows 4w .. 4w + 3. A row index is uniform over
  // a wavefront and a column index fixed per lane, so the staging's address arithmetic is a
  // few scalar operations per row and one multiply-add per lane (a flat halo index split
  // into row and column per slot cost most of the kernel's instructions).
  template <int p>
  struct OpTile
  {
    static constexpr int HX = 64, TX = HX - 2 * p, TY = 16, HY = TY + 2 * p;
    static constexpr int NWAVE = kThreads / 64, RO = TY / NWAVE, NR = (HY + NWAVE - 1) / NWAVE;
    static_assert(TY % NWAVE == 0, "owned rows per wavefront");
  };

  // The body of one operator block: block bid of the mesh's (XCD-ordered) tile blocks, vector v.
  template <typename V, int p, int KIND, int OP>
  __device__ __forceinline__ void
  operator_body(const Geo &G, const CgWork &W, const int it, const int tiles_x, const int tiles_y, const int bid,
                const int v, const V *x, const V *b, const V *r_in, V *y, V *y2)
  {
    using B                 = OpBands<OP>;
    using T                 = OpTile<p>;
    constexpr int  W_       = 2 * p + 1, HX = T::HX, HY = T::HY, TX = T::TX, TY = T::TY;
    constexpr int  RO       = T::RO, NR = T::NR, NWAVE = T::NWAVE;
    constexpr bool LAP      = OP == kOpLaplace;
    constexpr int  which[4] = {0, OP == kOpLaplace ? 2 : 1, 2, 3}; // band array of each slot
    __shared__ V      X[HY][HX];
    __shared__ double red[3][kThreads / 64];

    // No branch before the loads below (a branch on a kernel argument waits for it, and the
    // argument loads after it wait in turn: a chain of round trips): a padding block loads
    // the last tile's operand and leaves after the loads. The kernel arguments the loads
    // need are made available together first: their scalar loads are issued before the
    // statement (a zero into `early`, which the lane index takes on, so it stays ahead of
    // the loads; no memory effect, so uniform loads after it may still be scalar) and waited
    // for once, not each behind the previous one's wait at its first use.
    int early;
    asm("s_mov_b32 %0, 0"
        : "=s"(early)
        : "s"(tiles_x), "s"(tiles_y), "s"(it), "s"(G.P), "s"(G.np), "s"(G.stride), "s"(G.ax.periodic),
          "s"(G.ay.periodic), "s"(W.block[v]), "s"(x), "s"(r_in), "s"(G.dinv), "s"(W.state), "s"(W.state2),
          "s"(W.fold_rows), "s"(W.alpha_rows), "s"(W.partial), "s"(band_array(G.ax, which[0])),
          "s"(band_array(G.ay, which[0])), "s"(band_array(G.ax, which[B::NB - 1])),
          "s"(band_array(G.ay, which[B::NB - 1])), "s"(y), "s"(y2), "s"(b));
    const int  nt   = tiles_x * tiles_y;
    const int  nb8  = (nt + 7) & ~7;
    const int  tile = (bid & 7) * (nb8 >> 3) + (bid >> 3);
    const bool live = tile < nt;
    const int  tl   = live ? tile : nt - 1;
    const int  i0 = (tl % tiles_x) * TX, j0 = (tl / tiles_x) * TY;
    const int  P = G.P, tid = int(threadIdx.x);
    const int  lane = (tid & 63) + early, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const auto wall = [P](const int i, const int j) { return i == 0 || j == 0 || i == P || j == P; };
    // the staged node of halo coordinate i, -1 outside: walled, outside [0, P]; periodic, the
    // image in [0, P) of i in [-p, P + p] (a halo further out, past the last tile's nodes,
    // is never read and stages zero); selects, not branches
    const auto halo = [P](const int i, const int periodic) {
      int w = i < 0 ? i + P : i;
      w     = w >= P ? w - P : w;
      w     = w >= P ? w - P : w; // one cell of degree p: P = p
      const int per = i > P + p ? -1 : w;
      const int wal = (i < 0 || i > P) ? -1 : i;
      return periodic ? per : wal;
    };

    const int64_t vb = int64_t(W.block[v]) * G.stride;
    // Stage the operand (KIND 2: the new direction) of the halo'd block. A thread's NR rows
    // are loaded together, with the vector's state, before anything waits on them: one
    // memory round trip (this kernel's time at small meshes is its chain of round trips,
    // not its bytes). A converged vector's loads are in bounds and unused; so are those of a
    // slot outside the mesh, at an in-mesh address.
    // KIND 2: the stopping state after iteration it - 1, read, or (fold) formed here from
    // iteration it - 2's and the update stage's partial rows of iteration it - 1, summed as
    // the finalisation sums them (thread t takes rows t, t + 256, ..., then the block)
    constexpr int kFoldPer = kMaxCgFoldRows / kThreads; // update rows per thread
    const bool    fold     = KIND == 2 && W.fold_rows > 0 && it > 1;
    const double  st_word  = KIND == 2 ? state_word(states_of(W, fold ? it - 2 : it - 1) + v) : 0.;
    int           hx       = halo(i0 - p + lane, G.ax.periodic); // this lane's column
    const bool    col      = hx >= 0;
    hx                     = std::max(hx, 0);
    V             sx[NR], sr[NR];
    double        sd[NR];
    unsigned      in_mesh = 0; // per row slot: holds a node's value (the rest stage zero)
#pragma unroll
    for (int s = 0; s < NR; ++s)
      {
        const int ky = wave + NWAVE * s; // uniform
        int       hy = halo(j0 - p + std::min(ky, HY - 1), G.ay.periodic);
        // the Laplacian's wall entries are constrained to zero (read as zero by the cells)
        const bool ok = col && ky < HY && hy >= 0 && !(LAP && wall(hx, hy));
        in_mesh |= unsigned(ok) << s;
        hy                = std::max(hy, 0);
        const int64_t idx = int64_t(hy) * G.np + hx;
        // KIND 2 at it == 1 has no old direction: the buffer is read (it is there) and its
        // values unused
        sx[s] = x[vb + idx];
        sr[s] = V(0);
        sd[s] = 1.;
        if constexpr (KIND == 2)
          {
            sr[s] = r_in[vb + idx];
            sd[s] = G.dinv[idx];
          }
      }
    // the bands of this lane's owned column and of the wavefront's owned rows
    const int jr = wave * RO;
    const int gi = i0 + lane, gic = std::min(gi, P);
    V         xc[B::NB][W_], yc[RO][B::NB][W_];
#pragma unroll
    for (int o = 0; o < W_; ++o)
      {
#pragma unroll
        for (int t = 0; t < B::NB; ++t)
          xc[t][o] = V(band<p>(band_array(G.ax, which[t]), gic, o - p));
#pragma unroll
        for (int r = 0; r < RO; ++r)
#pragma unroll
          for (int t = 0; t < B::NB; ++t)
            yc[r][t][o] = V(band<p>(band_array(G.ay, which[t]), std::min(j0 + jr + r, P), o - p));
      }
    // the update stage's rows last: waiting for them is waiting for every load above
    double ua[kFoldPer], ub[kFoldPer];
#pragma unroll
    for (int j = 0; j < kFoldPer; ++j)
      ua[j] = ub[j] = 0.;
    if constexpr (KIND == 2)
#pragma unroll
      for (int j = 0; j < kFoldPer; ++j)
        {
          const int     bb  = tid + j * kThreads;
          const int     rc  = std::max(std::min(bb, W.fold_rows - 1), 0); // a row that is there
          const double *row = W.partial + size_t(W.alpha_rows + rc) * kPart + 2 * v;
          const double  a = row[0], b = row[1];
          ua[j]           = fold && bb < W.fold_rows ? a : 0.;
          ub[j]           = fold && bb < W.fold_rows ? b : 0.;
        }
    // rows and columns past the mesh hold zero bands
#pragma unroll
    for (int o = 0; o < W_; ++o)
      {
#pragma unroll
        for (int t = 0; t < B::NB; ++t)
          xc[t][o] = gi <= P ? xc[t][o] : V(0);
#pragma unroll
        for (int r = 0; r < RO; ++r)
#pragma unroll
          for (int t = 0; t < B::NB; ++t)
            yc[r][t][o] = j0 + jr + r <= P ? yc[r][t][o] : V(0);
      }
    if (!live)
      return;
    CgState st{};
    if constexpr (KIND == 2)
      {
        // reduced whether or not it is used, and the state formed in every thread (the loads
        // above stay unconditional and together that way)
        double u[2] = {0., 0.};
#pragma unroll
        for (int j = 0; j < kFoldPer; ++j)
          {
            u[0] += ua[j];
            u[1] += ub[j];
          }
        block_sum_all<2, kThreads>(u, red);
        st = state_from_lanes(st_word);
        if (fold)
          {
            st = advance_state(st, u[0], u[1], it - 1, W.max_its);
            if (tile == 0 && tid == 0) // one block per vector keeps it (the others read iteration it - 2's)
              states_of(W, it - 1)[v] = st;
          }
      }
    if (KIND == 2 && st.status != 0)
      {
        if (tid == 0)
          W.partial[size_t(tile) * kPart + 3 * v] = 0.;
        return;
      }
    const V beta = V((KIND == 2 && it > 1) ? st.beta : 0.);
#pragma unroll
    for (int s = 0; s < NR; ++s)
      {
        const int ky = wave + NWAVE * s;
        if (ky >= HY)
          continue;
        V val = V(0);
        if (in_mesh >> s & 1u)
          {
            val = sx[s];
            if constexpr (KIND == 2)
              {
                const V z = V(sd[s]) * sr[s];
                val       = it > 1 ? beta * sx[s] - z : -z;
              }
          }
        X[ky][lane] = val;
      }
    __syncthreads();

    // partial sums in scalars (an array indexed by a run-time vector number would live in scratch memory)
    const int ii    = std::min(lane, TX - 1); // lanes past the owned columns compute a copy, unused
    double    part0 = 0., part1 = 0., part2 = 0.;
    V         acc[RO];
#pragma unroll
    for (int r = 0; r < RO; ++r)
      acc[r] = V(0);
    const V c = V(W.vec[v].coef), ex = V(W.vec[v].e[0]), ey = V(W.vec[v].e[1]);
    if constexpr (OP == kOpLaplace) // Kx (x) My + Mx (x) Ky
      {
#pragma unroll
        for (int cc = 0; cc < W_; ++cc)
#pragma unroll
          for (int r = 0; r < RO; ++r)
            {
              V My = V(0), Ky = V(0);
#pragma unroll
              for (int o = 0; o < W_; ++o)
                {
                  const V xv = X[jr + r + o][ii + cc];
                  My += yc[r][B::M][o] * xv;
                  Ky += yc[r][B::K][o] * xv;
                }
              acc[r] += xc[B::K][cc] * My + xc[B::M][cc] * Ky;
            }
      }
    else
      {
        bool mass_only = true;
        if constexpr (OP == kOpStreaming)
          mass_only = c == V(0) || (ex == V(0) && ey == V(0));
        if (mass_only) // Mx (x) My
          {
#pragma unroll
            for (int cc = 0; cc < W_; ++cc)
#pragma unroll
              for (int r = 0; r < RO; ++r)
                {
                  V s = V(0);
#pragma unroll
                  for (int o = 0; o < W_; ++o)
                    s += yc[r][B::M][o] * X[jr + r + o][ii + cc];
                  acc[r] += xc[B::M][cc] * s;
                }
          }
        else if constexpr (OP == kOpStreaming)
          {
            const V kyy = c * ey * ey, kxx = c * ex * ex, kxy = c * ex * ey;
#pragma unroll
            for (int cc = 0; cc < W_; ++cc)
#pragma unroll
              for (int r = 0; r < RO; ++r)
                {
                  V My = V(0), Ky = V(0), Dy = V(0), DTy = V(0);
#pragma unroll
                  for (int o = 0; o < W_; ++o)
                    {
                      const V xv = X[jr + r + o][ii + cc];
                      My += yc[r][B::M][o] * xv;
                      Ky += yc[r][B::K][o] * xv;
                      Dy += yc[r][B::D][o] * xv;
                      DTy += yc[r][B::DT][o] * xv;
                    }
                  acc[r] += xc[B::M][cc] * (My + kyy * Ky) + xc[B::K][cc] * (kxx * My) + xc[B::DT][cc] * (kxy * Dy) +
                            xc[B::D][cc] * (kxy * DTy);
                }
          }
      }
#pragma unroll
    for (int r = 0; r < RO; ++r)
      {
        const int gj = j0 + jr + r;
        if (lane >= TX || gi > P || gj > P)
          continue;
        const int64_t idx   = node_index(G, gi, gj);
        const bool    slave = is_slave(G, gi, gj) || (LAP && wall(gi, gj));
        V             own   = X[jr + r + p][ii + p];
        V             a     = acc[r];
        if (slave) // constrained rows are the identity (MatrixFreeOperators::Base::vmult), on the
          {        // slave's own entry (the staged tile holds its master's, or zero on a wall)
            if constexpr (KIND == 2)
              own = it > 1 ? beta * x[vb + idx] - r_in[vb + idx] : -r_in[vb + idx]; // D^-1 = 1
            else
              own = x[vb + idx];
            a = own;
          }
        if constexpr (KIND == 0)
          y[vb + idx] = a;
        else if constexpr (KIND == 1)
          {
            const V bb  = b[vb + idx];
            const V rr  = a - bb;
            y[vb + idx] = rr;
            const double r8 = double(rr), b8 = double(bb);
            part0 += r8 * r8;
            part1 += b8 * b8;
            part2 += r8 * (mass_diag_inv(G, gi, gj) * r8);
          }
        else
          {
            y[vb + idx]  = own; // p_new
            y2[vb + idx] = a;   // q = A p_new
            part0 += double(own) * double(a);
          }
      }
    if constexpr (KIND == 1)
      {
        double pv[3] = {part0, part1, part2};
        block_sum<3>(pv, red);
        if (tid == 0)
          for (int k = 0; k < 3; ++k)
            W.partial[size_t(tile) * kPart + 3 * v + k] = pv[k];
      }
    else if constexpr (KIND == 2)
      {
        double pv[1] = {part0};
        block_sum<1>(pv, red);
        if (tid == 0)
          W.partial[size_t(tile) * kPart + 3 * v] = pv[0];
      }
  }

  template <typename V, int p, int KIND, int OP>
  __global__ void __launch_bounds__(kThreads)
  k_operator(const Geo G, const CgWork W, const int it, const int tiles_x, const int tiles_y, const V *x,
             const V *b, const V *r_in, V *y, V *y2)
  {
    operator_body<V, p, KIND, OP>(G, W, it, tiles_x, tiles_y, blockIdx.x, blockIdx.y, x, b, r_in, y, y2);
  }

  // the member of a batched launch that owns block b: the last whose first block is <= b
  template <typename M, typename F>
  __device__ __forceinline__ int
  member_of(const M *__restrict__ T, const int n, const int b, F first)
  {
    int lo = 0, hi = n - 1;
    while (lo < hi)
      {
        const int mid = (lo + hi + 1) >> 1;
        if (first(T[mid]) <= b)
          lo = mid;
        else
          hi = mid - 1;
      }
    return lo;
  }

  // The operator stage of many meshes' CGs in one launch: member m owns the blocks
  // [op_begin, op_begin + nb8) of the launch and runs exactly its own single-mesh blocks
  // (the CG start, KIND 1, on x, b, r; the direction, KIND 2, on the directions of
  // iteration it and q).
  template <int p, int KIND, int OP>
  __global__ void __launch_bounds__(kThreads)
  k_operator_batch(const CgMember *__restrict__ T, const int n, const int it)
  {
    const int       k = member_of(T, n, int(blockIdx.x), [](const CgMember &m) { return m.op_begin; });
    const CgMember &m = T[k];
    if (int(blockIdx.y) >= m.W.nv) // the launch has a row per population; a sweep may solve fewer
      return;
    const double   *x = KIND == 1 ? m.x : m.p[it % 2];
    double         *y = KIND == 1 ? m.r : m.p[(it + 1) % 2];
    operator_body<double, p, KIND, OP>(m.G, m.W, it, m.tiles_x, m.tiles_y, int(blockIdx.x) - m.op_begin,
                                               int(blockIdx.y), x, m.b, m.r, y, m.q);
  }

  // The scalars of one CG stage from the per-block partials, one block per vector and a
  // fixed summation order (a kernel boundary, not a device-wide fence, orders it after the
  // partials): thread t takes rows t, t + 256, ..., then the wavefront (shuffles), then the
  // wavefronts in order. STAGE 1: start (|r|, |b|, r.D^-1 r); 2: direction (alpha = rho /
  // p.q); 3: update (|r|, stopping test, rho, beta). The vectors' blocks read their partials
  // side by side on separate CUs (one block for all of them spent most of its time in the
  // partials' loads at a million nodes; 1024 threads in it measured no faster).
  constexpr int kFinThreads = kThreads;
  // nodes per thread and pass of the CG update (their loads issued together)
  constexpr int kUpdNodes = 2;

  template <int STAGE>
  __device__ __forceinline__ void
  finalize_body(const CgWork &W, const int v, const int nparts, const int it, const int row0)
  {
    constexpr int stride = STAGE == 3 ? 2 : 3, nk = STAGE == 1 ? 3 : STAGE == 2 ? 1 : 2;
    constexpr int NW = kFinThreads / 64;
    __shared__ double red[nk][NW];
    __shared__ double tot[nk];
    if (v >= W.nv)
      return;
    // the state and the partials loaded before the state's test (a converged vector's are in
    // bounds and unused); stage 3 turns iteration it - 1's state into iteration it's
    const CgState *in  = states_of(W, STAGE == 1 ? 0 : it - 1);
    CgState       *out = states_of(W, STAGE == 3 ? it : STAGE == 1 ? 0 : it - 1);
    CgState        st  = in[v];
    double         acc[nk];
#pragma unroll
    for (int k = 0; k < nk; ++k)
      acc[k] = 0.;
    for (int bb = threadIdx.x; bb < nparts; bb += kFinThreads)
#pragma unroll
      for (int k = 0; k < nk; ++k)
        acc[k] += W.partial[size_t(row0 + bb) * kPart + stride * v + k];
    if (STAGE != 1 && st.status != 0)
      {
        if (STAGE == 3 && threadIdx.x == 0) // carried to the next iteration's set
          out[v] = st;
        return;
      }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < nk; ++k)
      {
        const double t = wave_sum(acc[k]);
        if (lane == kWaveSumLane)
          red[k][wave] = t;
      }
    __syncthreads();
    if (threadIdx.x < nk)
      {
        double t = 0.;
#pragma unroll
        for (int w = 0; w < NW; ++w)
          t += red[threadIdx.x][w];
        tot[threadIdx.x] = t;
      }
    __syncthreads();
    if (threadIdx.x != 0)
      return;
    if constexpr (STAGE == 1)
      {
        st.res      = sqrt(tot[0]);
        st.tol      = W.tol_rel * sqrt(tot[1]);
        st.rho      = tot[2];
        st.rho_prev = 0.;
        st.its      = 0;
        // SolverControl::check(0, res): success first, then failure at the cap
        st.status = st.res <= st.tol ? 1 : (W.max_its <= 0 || isnan(st.res)) ? 2 : 0;
      }
    else if constexpr (STAGE == 2)
      st.alpha = st.rho / tot[0];
    else
      st = advance_state(st, tot[0], tot[1], it, W.max_its);
    out[v] = st;
  }

  template <int STAGE>
  __global__ void __launch_bounds__(kFinThreads)
  k_cg_finalize(const CgWork W, const int nparts, const int it, const int row0)
  {
    finalize_body<STAGE>(W, blockIdx.x, nparts, it, row0);
  }

  // every member's finalisation of one stage in one launch: block (member, vector); the
  // direction stage skips the members whose update kernel reduces its partials itself
  template <int STAGE>
  __global__ void __launch_bounds__(kFinThreads)
  k_cg_finalize_batch(const CgMember *__restrict__ T, const int it)
  {
    const CgMember &m = T[blockIdx.x];
    if (STAGE == 2 && m.W.alpha_rows > 0)
      return;
    finalize_body<STAGE>(m.W, blockIdx.y, STAGE == 3 ? m.upd_blocks : m.nt, it, STAGE == 3 ? m.W.alpha_rows : 0);
  }

  // x += alpha p; r += alpha q, with the partials of |r|^2 and r.D^-1 r (the stopping
  // test and beta follow in k_cg_finalize<3>). blockIdx.y is the vector: each block updates
  // one vector over the nodes its blockIdx.x strides through, so a batch of nv vectors has
  // nv times the wavefronts in flight, and each vector's sums keep the order they have with
  // one vector per launch.
  template <typename V>
  __device__ __forceinline__ void
  update_body(const Geo &G, const CgWork &W, V *__restrict__ x, V *__restrict__ r, const V *__restrict__ p,
              const V *__restrict__ q, const int v, const int blk, const int nblk, const int it)
  {
    __shared__ double red[2][kThreads / 64];
    // the kernel arguments the loads below need, loaded together (see operator_body)
    int early;
    asm("s_mov_b32 %0, 0"
        : "=s"(early)
        : "s"(W.nv), "s"(W.block[std::min(v, NM - 1)]), "s"(G.np), "s"(G.stride), "s"(G.dinv), "s"(W.state),
          "s"(W.state2), "s"(W.fold_rows), "s"(W.alpha_rows), "s"(W.partial), "s"(x), "s"(r), "s"(p), "s"(q),
          "s"(it), "s"(nblk));
    if (v >= W.nv)
      return;
    const int64_t n    = int64_t(G.np) * G.np;
    const int64_t off  = int64_t(W.block[v]) * G.stride;
    const int64_t step = int64_t(nblk) * kThreads * kUpdNodes;
    // the vector's state (across the lanes), the direction stage's partial rows for alpha
    // and this thread's first kUpdNodes nodes (base + u * kThreads) are loaded together: one
    // memory round trip; a converged vector's loads, and those of slots past the mesh (at its
    // last node), are in bounds and unused
    const double st_word = state_word(states_of(W, it - 1) + v); // after iteration it - 1
    int64_t      base    = blk * int64_t(kThreads) * kUpdNodes + int(threadIdx.x) + early;
    V            xv[kUpdNodes], pv[kUpdNodes], qv[kUpdNodes], rv[kUpdNodes];
    double       dv[kUpdNodes];
    const auto   load = [&](const int64_t b0) {
#pragma unroll
      for (int u = 0; u < kUpdNodes; ++u)
        {
          const int64_t kk = std::min(b0 + u * kThreads, n - 1);
          dv[u]            = G.dinv[kk];
          xv[u] = x[off + kk];
          pv[u] = p[off + kk];
          qv[u] = q[off + kk];
          rv[u] = r[off + kk];
        }
    };
    load(std::min(base, n - 1));
    // alpha = rho / p.q from the direction stage's partials, summed in a fixed order (thread t
    // takes rows t, t + 256, ..., then the block), so every block and thread holds the same
    // value, bit for bit, without a launch of its own; the first kAlphaPer rows of a thread
    // loaded with the rest
    constexpr int kAlphaPer = 2;
    double        pq[1]     = {0.};
    if (W.alpha_rows > 0)
      {
        double pr[kAlphaPer];
#pragma unroll
        for (int j = 0; j < kAlphaPer; ++j)
          {
            const int bb = int(threadIdx.x) + j * kThreads;
            const double a = W.partial[size_t(std::min(bb, W.alpha_rows - 1)) * kPart + 3 * v];
            pr[j]          = bb < W.alpha_rows ? a : 0.;
          }
#pragma unroll
        for (int j = 0; j < kAlphaPer; ++j)
          pq[0] += pr[j];
        for (int bb = int(threadIdx.x) + kAlphaPer * kThreads; bb < W.alpha_rows; bb += kThreads)
          pq[0] += W.partial[size_t(bb) * kPart + 3 * v];
        block_sum_all<1, kThreads>(pq, red);
      }
    const CgState st     = state_from_lanes(st_word);
    const int     status = st.status;
    const double  a      = W.alpha_rows > 0 ? st.rho / pq[0] : st.alpha;
    if (status != 0)
      return;
    const V alpha   = V(a);
    double  part[2] = {0., 0.};
    for (; base < n; base += step)
      {
        double r8[kUpdNodes], dk[kUpdNodes];
#pragma unroll
        for (int u = 0; u < kUpdNodes; ++u)
          {
            const int64_t kk = base + u * kThreads;
            const V       rr = rv[u] + alpha * qv[u];
            r8[u]            = double(rr);
            dk[u]            = dv[u];
            if (kk < n)
              {
                x[off + kk] = xv[u] + alpha * pv[u];
                r[off + kk] = rr;
              }
          }
        if (base + step < n) // the next pass's loads before this one's sums
          load(base + step);
#pragma unroll
        for (int u = 0; u < kUpdNodes; ++u)
          if (base + u * kThreads < n)
            {
              part[0] += r8[u] * r8[u];
              part[1] += r8[u] * (dk[u] * r8[u]);
            }
      }
    block_sum<2>(part, red);
    // after the direction stage's rows (other blocks may still be reading them for alpha)
    if (threadIdx.x == 0)
      {
        W.partial[size_t(W.alpha_rows + blk) * kPart + 2 * v]     = part[0];
        W.partial[size_t(W.alpha_rows + blk) * kPart + 2 * v + 1] = part[1];
      }
  }

  template <typename V>
  __global__ void __launch_bounds__(kThreads)
  k_cg_update(const Geo G, const CgWork W, V *__restrict__ x, V *__restrict__ r, const V *__restrict__ p,
              const V *__restrict__ q, const int it, const int nblk)
  {
    // nblk: gridDim.x as an argument (gridDim is read from the dispatch packet, a load of
    // its own)
    update_body<V>(G, W, x, r, p, q, blockIdx.y, blockIdx.x, nblk, it);
  }

  // every member's update in one launch: member m owns the blocks [upd_begin, upd_begin +
  // upd_blocks) of the launch, its single-mesh update grid
  __global__ void __launch_bounds__(kThreads)
  k_cg_update_batch(const CgMember *__restrict__ T, const int n, const int it)
  {
    const int       k = member_of(T, n, int(blockIdx.x), [](const CgMember &m) { return m.upd_begin; });
    const CgMember &m = T[k];
    update_body<double>(m.G, m.W, m.x, m.r, m.p[(it + 1) % 2], m.q, blockIdx.y, int(blockIdx.x) - m.upd_begin,
                        m.upd_blocks, it);
  }

  // The nodal stages of the batched consistent-mass step, elementwise over each member's
  // nodes (member m owns blocks [block_begin, block_begin + nblocks)): KIND 0 the
  // extrapolated start x = 2 x - x_prev, x_prev = x (streaming.h:76-80); 1 the increments
  // f_{a+1} += x_a (streaming.h:94-98); 2 Lee-Lin's predictor-corrector after the solves.
  template <int KIND>
  __global__ void __launch_bounds__(kThreads)
  k_nodal_batch(const NodalMember *__restrict__ T, const int n, const int parity)
  {
    const int          k = member_of(T, n, int(blockIdx.x), [](const NodalMember &m) { return m.block_begin; });
    const NodalMember &m = T[k];
    const Geo         &G = m.G;
    const int64_t      nn = int64_t(G.np) * G.np;
    double            *f  = m.f[(m.par + parity) & 1];
    for (int64_t q = int64_t(int(blockIdx.x) - m.block_begin) * kThreads + threadIdx.x; q < nn;
         q += int64_t(m.nblocks) * kThreads)
      {
        if constexpr (KIND == 0)
          {
#pragma unroll
            for (int a = 0; a < NM; ++a)
              if (m.mask & (1u << a))
                {
                  const double xa = m.x[a * G.stride + q], xp = m.x_prev[a * G.stride + q];
                  m.x[a * G.stride + q]      = -xp + 2. * xa;
                  m.x_prev[a * G.stride + q] = xa;
                }
          }
        else if constexpr (KIND == 1)
          {
#pragma unroll
            for (int a = 0; a < NM; ++a)
              if (m.mask & (1u << a))
                f[(a + 1) * G.stride + q] += m.x[a * G.stride + q];
          }
        else
          {
            const int i = int(q % G.np), j = int(q / G.np);
            double    v[Q], e[Q], x[NM];
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[a] = f[a * G.stride + q];
#pragma unroll
            for (int a = 0; a < NM; ++a)
              x[a] = m.x[a * G.stride + q];
            double     wx = 0., wy = 0.;
            const bool wall = wall_of(G, i, j, wx, wy);
            equilibrium_node(v, e, wall, wx, wy);
            predictor_corrector_node(v, e, x, m.theta, wall, wx, wy);
#pragma unroll
            for (int a = 0; a < Q; ++a)
              f[a * G.stride + q] = v[a];
          }
      }
  }

  // ------------------------------------------------------------------ stream function
  // The right-hand side of include/lbfem/io.h:61-81, (dphi/dx, v) - (dphi/dy, u), on the
  // tensor mesh: b = (Dx^T (x) My) v - (Mx (x) Dy^T) u, with the 1-D bands D_io = int phi_i
  // phi'_(i+o) and Dx^T its transpose; and the inverse of the Laplacian's diagonal
  // Kx_ii My_jj + Mx_ii Ky_jj (LaplaceOperator::compute_diagonal). Wall rows: b = 0, 1.
  template <int p>
  __global__ void __launch_bounds__(kThreads)
  k_psi_setup(const Geo G, const double *__restrict__ u, const double *__restrict__ v, double *__restrict__ b,
              double *__restrict__ dinv)
  {
    const int     P = G.P;
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        if (i == 0 || j == 0 || i == P || j == P)
          {
            b[k]    = 0.;
            dinv[k] = 1.;
            continue;
          }
        double s = 0.;
#pragma unroll
        for (int oy = -p; oy <= p; ++oy)
          {
            const int jj = j + oy;
            if (jj < 0 || jj > P)
              continue;
            const double my = band<p>(G.ay.M, j, oy), dty = band<p>(G.ay.DT, j, oy);
#pragma unroll
            for (int ox = -p; ox <= p; ++ox)
              {
                const int ii = i + ox;
                if (ii < 0 || ii > P)
                  continue;
                const int64_t q = node_index(G, ii, jj);
                s += band<p>(G.ax.DT, i, ox) * my * v[q] - band<p>(G.ax.M, i, ox) * dty * u[q];
              }
          }
        b[k]    = s;
        dinv[k] = 1. / (band<p>(G.ax.K, i, 0) * band<p>(G.ay.M, j, 0) + band<p>(G.ax.M, i, 0) * band<p>(G.ay.K, j, 0));
      }
  }

  // ------------------------------------------------------------------ diagnostics
  // T: double populations, or float deviations from the rest state (the mixed form);
  // u_exact (optional) scaled by ex_scale
  template <typename T>
  __global__ void __launch_bounds__(kThreads)
  k_diagnostics(const Geo G, const T *f, const double *u_exact, const double ex_scale, double *u_prev,
                double *partial, const int max_blocks)
  {
    __shared__ double red[6][kThreads / 64];
    const int64_t     n = int64_t(G.np) * G.np;
    double            s[6] = {0., 0., 0., 0., 0., 0.};
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        if (is_slave(G, i, j))
          continue;
        T v[Q];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = f[a * G.stride + k];
        double rho, ux, uy;
        if constexpr (sizeof(T) == 4)
          {
            double drho;
            moments_dev(v, drho, rho, ux, uy);
          }
        else
          moments(v, rho, ux, uy);
        const double w   = G.gen.ml ? G.gen.ml[k] : G.ax.ml[i] * G.ay.ml[j];
        const double exu = u_exact ? ex_scale * u_exact[2 * k] : 0.;
        const double exv = u_exact ? ex_scale * u_exact[2 * k + 1] : 0.;
        s[0] += w * ((ux - exu) * (ux - exu) + (uy - exv) * (uy - exv));
        s[1] += w * (exu * exu + exv * exv);
        s[2] += w * (ux * ux + uy * uy);
        s[3] += w * rho;
        s[4] += w;
        if (u_prev)
          {
            s[5] += w * hypot(ux - u_prev[2 * k], uy - u_prev[2 * k + 1]);
            u_prev[2 * k]     = ux;
            u_prev[2 * k + 1] = uy;
          }
      }
    block_sum<6>(s, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 6; ++k)
        partial[size_t(blockIdx.x) * kPart + k] = s[k];
  }

  __global__ void __launch_bounds__(kThreads)
  k_diagnostics_finalize(const double *partial, const int nparts, const int max_blocks, double *sums)
  {
    __shared__ double red[6][kThreads / 64];
    (void)max_blocks;
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    for (int bb = threadIdx.x; bb < nparts; bb += kThreads)
#pragma unroll
      for (int k = 0; k < 6; ++k)
        acc[k] += partial[size_t(bb) * kPart + k];
    block_sum<6>(acc, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 6; ++k)
        sums[k] = acc[k];
  }

  // ------------------------------------------------------------------ velocity-change monitor
  __global__ void __launch_bounds__(kThreads)
  k_monitor(const MonitorEntry *__restrict__ E, const int n_members, double *__restrict__ partial)
  {
    __shared__ double red[3][kThreads / 64];
    __shared__ double mx[kThreads / 64];
    const int         b  = blockIdx.x;
    int               lo = 0, hi = n_members - 1; // the member: last block_begin <= b
    while (lo < hi)
      {
        const int mid = (lo + hi + 1) >> 1;
        if (E[mid].block_begin <= b)
          lo = mid;
        else
          hi = mid - 1;
      }
    const MonitorEntry &e = E[lo];
    const Geo          &G = e.G;
    const int64_t       n = int64_t(G.np) * G.np;
    double              dmax = 0., s[3] = {0., 0., 0.};
    for (int64_t k = int64_t(b - e.block_begin) * kThreads + threadIdx.x; k < n; k += int64_t(e.nblocks) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        if (is_slave(G, i, j))
          continue;
        double rho, ux, uy;
        if (e.fp32)
          {
            const float *f = static_cast<const float *>(e.f);
            float        v[Q];
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[a] = f[a * G.stride + k];
            double drho;
            moments_dev(v, drho, rho, ux, uy);
          }
        else
          {
            const double *f = static_cast<const double *>(e.f);
            double        v[Q];
#pragma unroll
            for (int a = 0; a < Q; ++a)
              v[a] = f[a * G.stride + k];
            moments(v, rho, ux, uy);
          }
        if (!(isfinite(rho) && isfinite(ux) && isfinite(uy)))
          {
            s[2] += 1.;
            continue;
          }
        if (!e.init)
          {
            const double w  = G.gen.ml ? G.gen.ml[k] : G.ax.ml[i] * G.ay.ml[j];
            const double du = hypot(ux - e.u_prev[2 * k], uy - e.u_prev[2 * k + 1]);
            dmax            = fmax(dmax, du);
            s[0] += w * du;
            s[1] += w;
          }
        e.u_prev[2 * k]     = ux;
        e.u_prev[2 * k + 1] = uy;
      }
    // the maximum (order-free), then the sums in a fixed order
    for (int off = 32; off > 0; off >>= 1)
      dmax = fmax(dmax, __shfl_down(dmax, off, 64));
    if ((threadIdx.x & 63) == 0)
      mx[threadIdx.x >> 6] = dmax;
    block_sum<3>(s, red);
    if (threadIdx.x == 0)
      {
        double m = 0.;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w)
          m = fmax(m, mx[w]);
        partial[size_t(b) * kMonitorPartials + 0] = m;
        partial[size_t(b) * kMonitorPartials + 1] = s[0];
        partial[size_t(b) * kMonitorPartials + 2] = s[1];
        partial[size_t(b) * kMonitorPartials + 3] = s[2];
      }
  }

  // one block per member: its blocks' partials in block order
  __global__ void __launch_bounds__(64)
  k_monitor_finalize(const MonitorEntry *__restrict__ E, const double *__restrict__ partial,
                     double *__restrict__ results)
  {
    const MonitorEntry &e = E[blockIdx.x];
    double              m = 0., s[3] = {0., 0., 0.};
    for (int b = threadIdx.x; b < e.nblocks; b += 64)
      {
        const double *q = partial + size_t(e.block_begin + b) * kMonitorPartials;
        m               = fmax(m, q[0]);
        s[0] += q[1];
        s[1] += q[2];
        s[2] += q[3];
      }
    for (int off = 32; off > 0; off >>= 1)
      {
        m = fmax(m, __shfl_down(m, off, 64));
#pragma unroll
        for (int k = 0; k < 3; ++k)
          s[k] += __shfl_down(s[k], off, 64);
      }
    if (threadIdx.x == 0)
      {
        double *r = results + size_t(blockIdx.x) * kMonitorPartials;
        r[0]      = e.init ? HUGE_VAL : m;
        r[1]      = e.init ? HUGE_VAL : s[0];
        r[2]      = s[1];
        r[3]      = s[2];
      }
  }

  hipError_t
  monitor_launch(const MonitorEntry *table, const int n_members, const int total_blocks, double *partial,
                 double *results, hipStream_t s)
  {
    if (n_members <= 0)
      return hipSuccess;
    hipLaunchKernelGGL(k_monitor, dim3(total_blocks), dim3(kThreads), 0, s, table, n_members, partial);
    hipLaunchKernelGGL(k_monitor_finalize, dim3(n_members), dim3(64), 0, s, table, partial, results);
    return hipGetLastError();
  }

  // ------------------------------------------------------------------ launchers
  static int
  nodal_blocks(const Geo &G)
  {
    const int64_t n = int64_t(G.np) * G.np;
    const int64_t b = (n + kThreads - 1) / kThreads;
    return int(b < 4096 ? b : 4096);
  }

  hipError_t
  collide_launch(const Geo &G, double *g, const double omega, hipStream_t s)
  {
    hipLaunchKernelGGL(k_collide, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, g, omega);
    return hipGetLastError();
  }
  hipError_t
  equilibrium_launch(const Geo &G, const double *f, double *feq, hipStream_t s)
  {
    hipLaunchKernelGGL(k_equilibrium, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, f, feq);
    return hipGetLastError();
  }
  hipError_t
  predictor_corrector_launch(const Geo &G, double *f, const double *feq, const double *incr, const double theta,
                             hipStream_t s)
  {
    hipLaunchKernelGGL(k_predictor_corrector, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, f, feq, incr, theta);
    return hipGetLastError();
  }
  hipError_t
  pc_from_f_launch(const Geo &G, double *f, const double *incr, const double theta, hipStream_t s)
  {
    hipLaunchKernelGGL(k_pc_from_f, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, f, incr, theta);
    return hipGetLastError();
  }
  hipError_t
  add_increments_launch(const Geo &G, double *f, const double *x, const unsigned mask, hipStream_t s)
  {
    hipLaunchKernelGGL(k_add_increments, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, f, x, mask);
    return hipGetLastError();
  }
  hipError_t
  extrapolate_launch(const Geo &G, double *x, double *x_prev, const unsigned mask, hipStream_t s)
  {
    hipLaunchKernelGGL(k_extrapolate<double>, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, x, x_prev, mask);
    return hipGetLastError();
  }
  hipError_t
  psi_setup_launch(const Geo &G, const double *u, const double *v, double *b, double *dinv, hipStream_t s)
  {
    switch (G.p)
      {
        case 1:
          hipLaunchKernelGGL(k_psi_setup<1>, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, u, v, b, dinv);
          break;
        case 2:
          hipLaunchKernelGGL(k_psi_setup<2>, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, u, v, b, dinv);
          break;
        case 3:
          hipLaunchKernelGGL(k_psi_setup<3>, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, u, v, b, dinv);
          break;
        default:
          return hipErrorInvalidValue;
      }
    return hipGetLastError();
  }
  hipError_t
  lumped_solve_launch(const Geo &G, double *x, const double *r, const int nv, hipStream_t s)
  {
    hipLaunchKernelGGL(k_lumped_solve, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, x, r, nv);
    return hipGetLastError();
  }
  hipError_t
  richardson_update_launch(const Geo &G, double *x, const double *r, const double *Ax, const int nv, hipStream_t s)
  {
    hipLaunchKernelGGL(k_richardson_update, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, x, r, Ax, nv);
    return hipGetLastError();
  }

  template <int KIND, typename V>
  static hipError_t
  operator_launch(const Geo &G, const CgWork &W, const int it, const V *x, const V *b, const V *r, V *y, V *y2,
                  hipStream_t s)
  {
    if (G.gen.vert != nullptr) // general quadrilaterals: the cell-wise operator (FP64 only)
      {
        if constexpr (std::is_same_v<V, double>)
          {
            const hipError_t e = quad_operator_launch<KIND>(G, W, it, x, b, r, y, y2, s);
            if (e != hipSuccess)
              return e;
            if (KIND != 0 && !(KIND == 2 && W.alpha_rows > 0))
              hipLaunchKernelGGL((k_cg_finalize<KIND>), dim3(std::max(W.nv, 1)), dim3(kFinThreads), 0, s, W, quad_tile_count(G), it,
                                 0);
            return hipGetLastError();
          }
        else
          return hipErrorInvalidValue;
      }
    int tx = 0, ty = 0;
    cg_operator_tiles(G, tx, ty);
    const int nt = tx * ty, nb8 = (nt + 7) & ~7;
    if (KIND != 0 && nt > W.max_blocks)
      return hipErrorInvalidValue;
    // the band set: the Laplacian's, or the streaming matrix's when a vector has a K_e term
    int op = W.laplace ? int(kOpLaplace) : int(kOpMass);
    for (int v = 0; v < W.nv && !W.laplace; ++v)
      if (W.vec[v].coef != 0. && (W.vec[v].e[0] != 0. || W.vec[v].e[1] != 0.))
        op = kOpStreaming;
    const dim3 grid(nb8, W.nv), block(kThreads);
    const auto go = [&](auto pc, auto oc) {
      constexpr int P_ = decltype(pc)::value, OP_ = decltype(oc)::value;
      hipLaunchKernelGGL((k_operator<V, P_, KIND, OP_>), grid, block, 0, s, G, W, it, tx, ty, x, b, r, y, y2);
    };
    const auto by_op = [&](auto pc) {
      if (op == kOpMass)
        go(pc, std::integral_constant<int, kOpMass>{});
      else if (op == kOpStreaming)
        go(pc, std::integral_constant<int, kOpStreaming>{});
      else if constexpr (std::is_same_v<V, double>)
        go(pc, std::integral_constant<int, kOpLaplace>{});
      else
        return false; // the stream function is FP64 only
      return true;
    };
    bool ok = false;
    switch (G.p)
      {
        case 1:
          ok = by_op(std::integral_constant<int, 1>{});
          break;
        case 2:
          ok = by_op(std::integral_constant<int, 2>{});
          break;
        case 3:
          ok = by_op(std::integral_constant<int, 3>{});
          break;
        default:
          break;
      }
    if (!ok)
      return hipErrorInvalidValue;
    if (KIND != 0 && !(KIND == 2 && W.alpha_rows > 0))
      hipLaunchKernelGGL((k_cg_finalize<KIND>), dim3(std::max(W.nv, 1)), dim3(kFinThreads), 0, s, W, nt, it, 0);
    return hipGetLastError();
  }

  static_assert(OpTile<1>::TX == 62 && OpTile<3>::TX == 58 && OpTile<2>::TY == 16, "cg_operator_tiles");

  void
  cg_operator_tiles(const Geo &G, int &tiles_x, int &tiles_y)
  {
    const int tx = 64 - 2 * G.p, ty = 16; // OpTile<p>::TX, TY
    tiles_x      = (G.P + 1 + tx - 1) / tx;
    tiles_y      = (G.P + 1 + ty - 1) / ty;
  }

  int
  cg_direction_rows(const Geo &G)
  {
    if (G.gen.vert != nullptr)
      return quad_tile_count(G);
    int tx = 0, ty = 0;
    cg_operator_tiles(G, tx, ty);
    return tx * ty;
  }

  int
  cg_update_blocks(const Geo &G, const int max_blocks, const int alpha_rows)
  {
    const int64_t n    = int64_t(G.np) * G.np;
    const int64_t need = (n + int64_t(kThreads) * kUpdNodes - 1) / (int64_t(kThreads) * kUpdNodes);
    return int(std::min<int64_t>(need, std::min(max_blocks - alpha_rows, 1024)));
  }

  int
  nodal_batch_blocks(const Geo &G)
  {
    return nodal_blocks(G);
  }

  template <typename V>
  static hipError_t
  update_launch(const Geo &G, const CgWork &W, const int it, V *x, V *r, const V *p, const V *q, hipStream_t s)
  {
    const int nb = cg_update_blocks(G, W.max_blocks, W.alpha_rows);
    hipLaunchKernelGGL(k_cg_update<V>, dim3(nb, std::max(W.nv, 1)), dim3(kThreads), 0, s, G, W, x, r, p, q, it, nb);
    if (W.fold_rows == 0) // else the next direction stage folds it in
      hipLaunchKernelGGL((k_cg_finalize<3>), dim3(std::max(W.nv, 1)), dim3(kFinThreads), 0, s, W, nb, it,
                         W.alpha_rows);
    return hipGetLastError();
  }

  hipError_t
  cg_update_finalize_launch(const Geo &G, const CgWork &W, const int it, hipStream_t s)
  {
    const int nb = cg_update_blocks(G, W.max_blocks, W.alpha_rows);
    hipLaunchKernelGGL((k_cg_finalize<3>), dim3(std::max(W.nv, 1)), dim3(kFinThreads), 0, s, W, nb, it, W.alpha_rows);
    return hipGetLastError();
  }

  // ------------------------------------------------------------------ batched CG launchers
  template <int KIND>
  static hipError_t
  operator_batch_launch(const int p, const int op, const CgMember *T, const int n, const int op_blocks, const int it,
                        hipStream_t s)
  {
    const dim3 grid(op_blocks, NM), block(kThreads);
    const auto go = [&](auto pc, auto oc) {
      constexpr int P_ = decltype(pc)::value, OP_ = decltype(oc)::value;
      hipLaunchKernelGGL((k_operator_batch<P_, KIND, OP_>), grid, block, 0, s, T, n, it);
    };
    const auto by_op = [&](auto pc) {
      if (op == kOpMass)
        go(pc, std::integral_constant<int, kOpMass>{});
      else
        go(pc, std::integral_constant<int, kOpStreaming>{});
    };
    switch (p)
      {
        case 1:
          by_op(std::integral_constant<int, 1>{});
          break;
        case 2:
          by_op(std::integral_constant<int, 2>{});
          break;
        case 3:
          by_op(std::integral_constant<int, 3>{});
          break;
        default:
          return hipErrorInvalidValue;
      }
    return hipGetLastError();
  }

  hipError_t
  cg_batch_start_launch(const int p, const int op, const CgMember *T, const int n, const int op_blocks, hipStream_t s)
  {
    if (n <= 0)
      return hipSuccess;
    if (const hipError_t e = operator_batch_launch<1>(p, op, T, n, op_blocks, 0, s); e != hipSuccess)
      return e;
    hipLaunchKernelGGL((k_cg_finalize_batch<1>), dim3(n, NM), dim3(kFinThreads), 0, s, T, 0);
    return hipGetLastError();
  }

  hipError_t
  cg_batch_direction_launch(const int p, const int op, const CgMember *T, const int n, const int op_blocks,
                            const int it, hipStream_t s)
  {
    if (n <= 0)
      return hipSuccess;
    if (const hipError_t e = operator_batch_launch<2>(p, op, T, n, op_blocks, it, s); e != hipSuccess)
      return e;
    hipLaunchKernelGGL((k_cg_finalize_batch<2>), dim3(n, NM), dim3(kFinThreads), 0, s, T, it);
    return hipGetLastError();
  }

  hipError_t
  cg_batch_update_launch(const CgMember *T, const int n, const int upd_blocks, const int it, hipStream_t s)
  {
    if (n <= 0)
      return hipSuccess;
    hipLaunchKernelGGL(k_cg_update_batch, dim3(upd_blocks, NM), dim3(kThreads), 0, s, T, n, it);
    hipLaunchKernelGGL((k_cg_finalize_batch<3>), dim3(n, NM), dim3(kFinThreads), 0, s, T, it);
    return hipGetLastError();
  }

  hipError_t
  nodal_batch_launch(const int kind, const NodalMember *T, const int n, const int blocks, const int parity,
                     hipStream_t s)
  {
    if (n <= 0)
      return hipSuccess;
    switch (kind)
      {
        case 0:
          hipLaunchKernelGGL(k_nodal_batch<0>, dim3(blocks), dim3(kThreads), 0, s, T, n, parity);
          break;
        case 1:
          hipLaunchKernelGGL(k_nodal_batch<1>, dim3(blocks), dim3(kThreads), 0, s, T, n, parity);
          break;
        case 2:
          hipLaunchKernelGGL(k_nodal_batch<2>, dim3(blocks), dim3(kThreads), 0, s, T, n, parity);
          break;
        default:
          return hipErrorInvalidValue;
      }
    return hipGetLastError();
  }

  hipError_t
  vmult_launch(const Geo &G, const CgWork &W, const double *x, double *y, hipStream_t s)
  {
    return operator_launch<0, double>(G, W, 0, x, nullptr, nullptr, y, nullptr, s);
  }
  hipError_t
  cg_start_launch(const Geo &G, const CgWork &W, const double *x, const double *b, double *r, hipStream_t s)
  {
    return operator_launch<1, double>(G, W, 0, x, b, nullptr, r, nullptr, s);
  }
  hipError_t
  cg_direction_launch(const Geo &G, const CgWork &W, const int it, const double *p_old, const double *r,
                      double *p_new, double *q, hipStream_t s)
  {
    return operator_launch<2, double>(G, W, it, p_old, nullptr, r, p_new, q, s);
  }
  hipError_t
  cg_update_launch(const Geo &G, const CgWork &W, const int it, double *x, double *r, const double *p, const double *q,
                   hipStream_t s)
  {
    return update_launch<double>(G, W, it, x, r, p, q, s);
  }
  hipError_t
  cg_start_launch(const Geo &G, const CgWork &W, const float *x, const float *b, float *r, hipStream_t s)
  {
    return operator_launch<1, float>(G, W, 0, x, b, nullptr, r, nullptr, s);
  }
  hipError_t
  cg_direction_launch(const Geo &G, const CgWork &W, const int it, const float *p_old, const float *r,
                      float *p_new, float *q, hipStream_t s)
  {
    return operator_launch<2, float>(G, W, it, p_old, nullptr, r, p_new, q, s);
  }
  hipError_t
  cg_update_launch(const Geo &G, const CgWork &W, const int it, float *x, float *r, const float *p, const float *q,
                   hipStream_t s)
  {
    return update_launch<float>(G, W, it, x, r, p, q, s);
  }

  // ------------------------------------------------------------------ mixed-form helpers
  // Bardow: g_{a+1} += x_a on FP32 deviations from FP32 increments (the sum in FP64, rounded once)
  __global__ void __launch_bounds__(kThreads)
  k_add_increments_mixed(const Geo G, float *g, const float *x, const unsigned mask)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
#pragma unroll
      for (int a = 0; a < NM; ++a)
        if (mask & (1u << a))
          g[(a + 1) * G.stride + k] = float(double(g[(a + 1) * G.stride + k]) + double(x[a * G.stride + k]));
  }

  // Lee-Lin after the consistent-mass solve, on FP32 deviations: f <- PC(f, feq(f), x)
  __global__ void __launch_bounds__(kThreads)
  k_pc_from_f_mixed(const Geo G, float *f, const float *incr, const double theta)
  {
    const int64_t n = int64_t(G.np) * G.np;
    for (int64_t k = blockIdx.x * int64_t(kThreads) + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
      {
        const int i = int(k % G.np), j = int(k / G.np);
        float     v[Q], fe[Q], x[NM];
#pragma unroll
        for (int a = 0; a < Q; ++a)
          v[a] = f[a * G.stride + k];
#pragma unroll
        for (int a = 0; a < NM; ++a)
          x[a] = incr[a * G.stride + k];
        double     wx = 0., wy = 0.;
        const bool wall = wall_of(G, i, j, wx, wy);
        equilibrium_node(v, fe, wall, wx, wy);
        predictor_corrector_node(v, fe, x, theta, wall, wx, wy);
#pragma unroll
        for (int a = 0; a < Q; ++a)
          f[a * G.stride + k] = v[a];
      }
  }

  hipError_t
  extrapolate_launch(const Geo &G, float *x, float *x_prev, const unsigned mask, hipStream_t s)
  {
    hipLaunchKernelGGL(k_extrapolate<float>, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, x, x_prev, mask);
    return hipGetLastError();
  }
  hipError_t
  add_increments_launch(const Geo &G, float *g, const float *x, const unsigned mask, hipStream_t s)
  {
    hipLaunchKernelGGL(k_add_increments_mixed, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, g, x, mask);
    return hipGetLastError();
  }
  hipError_t
  pc_from_f_launch(const Geo &G, float *f, const float *incr, const double theta, hipStream_t s)
  {
    hipLaunchKernelGGL(k_pc_from_f_mixed, dim3(nodal_blocks(G)), dim3(kThreads), 0, s, G, f, incr, theta);
    return hipGetLastError();
  }
  hipError_t
  diagnostics_launch(const Geo &G, const double *f, const double *u_exact, double *u_prev, double *sums,
                     double *partial, const int max_blocks, hipStream_t s)
  {
    int nb = std::min(nodal_blocks(G), std::min(max_blocks, 1024));
    hipLaunchKernelGGL(k_diagnostics<double>, dim3(nb), dim3(kThreads), 0, s, G, f, u_exact, 1., u_prev, partial,
                       max_blocks);
    hipLaunchKernelGGL(k_diagnostics_finalize, dim3(1), dim3(kThreads), 0, s, partial, nb, max_blocks, sums);
    return hipGetLastError();
  }

  template <typename T>
  hipError_t
  diagnostics_scaled_launch(const Geo &G, const T *f, const double *u_exact, const double ex_scale, double *u_prev,
                            double *sums, double *partial, const int max_blocks, hipStream_t s)
  {
    int nb = std::min(nodal_blocks(G), std::min(max_blocks, 1024));
    hipLaunchKernelGGL(k_diagnostics<T>, dim3(nb), dim3(kThreads), 0, s, G, f, u_exact, ex_scale, u_prev, partial,
                       max_blocks);
    hipLaunchKernelGGL(k_diagnostics_finalize, dim3(1), dim3(kThreads), 0, s, partial, nb, max_blocks, sums);
    return hipGetLastError();
  }
  template hipError_t diagnostics_scaled_launch<double>(const Geo &, const double *, const double *, double,
                                                        double *, double *, double *, int, hipStream_t);
  template hipError_t diagnostics_scaled_launch<float>(const Geo &, const float *, const double *, double, double *,
                                                       double *, double *, int, hipStream_t);
  // ---------------------------------------------------------------- HBM probe
  // A plain copy: each block copies one contiguous chunk, four 16 B non-temporal
  // loads in flight per lane per pass, non-temporal stores (the fastest of the copy
  // shapes measured on this device: 6.2-6.3 TB/s read + write over 1 GiB, against
  // 4.5 for a grid-strided copy with cached stores): the achievable-bandwidth
  // reference and the known-size kernel that scales the byte counters of the
  // profiler.
  typedef double v2d __attribute__((ext_vector_type(2)));
  constexpr int kCopyBlocks = 32768;

  __global__ void __launch_bounds__(kThreads)
  k_stream_copy(const v2d *__restrict__ src, v2d *__restrict__ dst, const size_t n, const size_t chunk)
  {
    constexpr int U    = 4;
    const size_t  base = size_t(blockIdx.x) * chunk;
    const size_t  end  = base + chunk < n ? base + chunk : n;
    for (size_t i = base + threadIdx.x; i < end; i += U * kThreads)
      {
        v2d v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (i + u * kThreads < end)
            v[u] = __builtin_nontemporal_load(src + i + u * kThreads);
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (i + u * kThreads < end)
            __builtin_nontemporal_store(v[u], dst + i + u * kThreads);
      }
  }

  hipError_t
  stream_copy_launch(void *dst, const void *src, const size_t bytes, hipStream_t s)
  {
    const size_t n = bytes / sizeof(v2d);
    if (n == 0)
      return hipSuccess;
    const size_t chunk = (n + kCopyBlocks - 1) / kCopyBlocks;
    const size_t nb    = (n + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_stream_copy, dim3(unsigned(nb)), dim3(kThreads), 0, s, static_cast<const v2d *>(src),
                       static_cast<v2d *>(dst), n, chunk);
    return hipGetLastError();
  }
} // namespace lbfem_amd
