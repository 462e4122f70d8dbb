// rism_fft.h -- LDS-staged mixed-radix (2,3,4,5,8) Stockham FFT of columns,
// the building block of the axis passes in rism_kernels.hip.
//
// One column (a line of L complex values along one grid axis) is transformed
// by T threads of a workgroup.  Stage s with radix R performs L/R radix-R
// butterflies; butterfly j reads elements j + r*L/R and writes elements
// (j/Ns)*Ns*R + j%Ns + r*Ns, Ns being the product of the earlier radices
// (the auto-sorting Stockham form, so no bit reversal pass is needed).
// The first stage can read straight from global memory through a Load
// functor and the last stage can write straight to global memory through a
// Store functor; both touch element i at (i, column), so with the column
// index fastest across lanes every global access is coalesced.  Intermediate
// stages exchange through LDS, padded one complex every 8 to break the
// stride-R bank conflicts of the butterfly outputs.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace rism {

// Complex value of scalar type Re (double: the FP64 path; float: the FP32
// bulk of the mixed-precision path), aligned to its size so that LDS and
// global accesses are single 16-byte (8-byte) vector operations.
template <class Re>
struct alignas(2 * sizeof(Re)) cx {
  Re x, y;
};
using cplx = cx<double>;
using cplxf = cx<float>;

template <class Re>
struct ident {
  using type = Re;
};

template <class Re>
__device__ __forceinline__ cx<Re> mkc(Re a, Re b) {
  return cx<Re>{a, b};
}
__device__ __forceinline__ cplx mk(double a, double b) { return cplx{a, b}; }
template <class Re>
__device__ __forceinline__ cx<Re> operator+(cx<Re> a, cx<Re> b) {
  return cx<Re>{a.x + b.x, a.y + b.y};
}
template <class Re>
__device__ __forceinline__ cx<Re> operator-(cx<Re> a, cx<Re> b) {
  return cx<Re>{a.x - b.x, a.y - b.y};
}
template <class Re>
__device__ __forceinline__ cx<Re> operator*(typename ident<Re>::type s, cx<Re> a) {
  return cx<Re>{s * a.x, s * a.y};
}
template <class Re>
__device__ __forceinline__ cx<Re> cmul(cx<Re> a, cx<Re> b) {
  return cx<Re>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <class Re>
__device__ __forceinline__ cx<Re> cmulc(cx<Re> a, cx<Re> b) {
  return cx<Re>{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <class Re>
__device__ __forceinline__ cx<Re> conj(cx<Re> a) {
  return cx<Re>{a.x, -a.y};
}
// multiply by S*i, S = +1 or -1
template <int S, class Re>
__device__ __forceinline__ cx<Re> mul_si(cx<Re> a) {
  return S > 0 ? cx<Re>{-a.y, a.x} : cx<Re>{a.y, -a.x};
}

// ---- radix-R DFT in registers; S = -1 forward, +1 inverse ----------------
template <int R, int S>
struct Dft;

template <int S>
struct Dft<2, S> {
  template <class Re>
  __device__ __forceinline__ static void run(cx<Re> *v) {
    cx<Re> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <int S>
struct Dft<4, S> {
  template <class Re>
  __device__ __forceinline__ static void run(cx<Re> *v) {
    cx<Re> s02 = v[0] + v[2], d02 = v[0] - v[2];
    cx<Re> s13 = v[1] + v[3], d13 = mul_si<S>(v[1] - v[3]);
    v[0] = s02 + s13;
    v[2] = s02 - s13;
    v[1] = d02 + d13;
    v[3] = d02 - d13;
  }
};

template <int S>
struct Dft<8, S> {
  template <class Re>
  __device__ __forceinline__ static void run(cx<Re> *v) {
    constexpr Re h = (Re)0.70710678118654752440;
    cx<Re> e[4] = {v[0], v[2], v[4], v[6]};
    cx<Re> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, S>::run(e);
    Dft<4, S>::run(o);
    // w8^k = exp(S 2 pi i k / 8)
    cx<Re> o1 = cmul(o[1], cx<Re>{h, S * h});
    cx<Re> o2 = mul_si<S>(o[2]);
    cx<Re> o3 = cmul(o[3], cx<Re>{-h, S * h});
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + o1;
    v[5] = e[1] - o1;
    v[2] = e[2] + o2;
    v[6] = e[2] - o2;
    v[3] = e[3] + o3;
    v[7] = e[3] - o3;
  }
};

template <int S>
struct Dft<3, S> {
  template <class Re>
  __device__ __forceinline__ static void run(cx<Re> *v) {
    constexpr Re s3 = (Re)0.86602540378443864676;  // sqrt(3)/2
    cx<Re> t1 = v[1] + v[2];
    cx<Re> t2 = v[0] - (Re)0.5 * t1;
    cx<Re> t3 = mul_si<S>(s3 * (v[1] - v[2]));
    v[0] = v[0] + t1;
    v[1] = t2 + t3;
    v[2] = t2 - t3;
  }
};

template <int S>
struct Dft<5, S> {
  template <class Re>
  __device__ __forceinline__ static void run(cx<Re> *v) {
    constexpr Re c1 = (Re)0.30901699437494742410;   // cos(2pi/5)
    constexpr Re c2 = (Re)-0.80901699437494742410;  // cos(4pi/5)
    constexpr Re s1 = (Re)0.95105651629515357212;   // sin(2pi/5)
    constexpr Re s2 = (Re)0.58778525229247312917;   // sin(4pi/5)
    cx<Re> t1 = v[1] + v[4], t2 = v[2] + v[3];
    cx<Re> t3 = v[1] - v[4], t4 = v[2] - v[3];
    cx<Re> b1 = v[0] + c1 * t1 + c2 * t2;
    cx<Re> b2 = v[0] + c2 * t1 + c1 * t2;
    cx<Re> d1 = mul_si<S>(s1 * t3 + s2 * t4);
    cx<Re> d2 = mul_si<S>(s2 * t3 - s1 * t4);
    v[0] = v[0] + t1 + t2;
    v[1] = b1 + d1;
    v[4] = b1 - d1;
    v[2] = b2 + d2;
    v[3] = b2 - d2;
  }
};

// ---- compile-time factorisation ------------------------------------------
struct RadixList {
  int n;
  int r[16];
};

// maxr = 8: as many radix-8 stages as possible; maxr = 4: radix-4 stages
// (twice the threads per column, smaller butterflies)
constexpr RadixList make_radices(int L, int maxr = 8) {
  RadixList rl{0, {}};
  int m = L;
  int n8 = 0, n4 = 0, n2 = 0;
  while (m % 2 == 0) { ++n2; m /= 2; }
  int twos = 0;
  if (maxr == 4) {
    n4 = n2 / 2;
    twos = n2 % 2;
  } else {
    // powers of two: as many radix-8 as possible, remainder as 4 / 4+? / 2
    n8 = n2 / 3;
    int rem = n2 % 3;
    if (rem == 1) {
      if (n8 > 0) { n8 -= 1; n4 = 2; } else { n2 = 1; }
    } else if (rem == 2) {
      n4 = 1;
    }
    twos = (rem == 1 && n4 == 0) ? 1 : 0;
  }
  for (int i = 0; i < n8; ++i) rl.r[rl.n++] = 8;
  for (int i = 0; i < n4; ++i) rl.r[rl.n++] = 4;
  for (int i = 0; i < twos; ++i) rl.r[rl.n++] = 2;
  while (m % 5 == 0) { rl.r[rl.n++] = 5; m /= 5; }
  while (m % 3 == 0) { rl.r[rl.n++] = 3; m /= 3; }
  if (m != 1) rl.n = -1;  // unsupported prime factor
  return rl;
}

// whether the last stage of the plan of length L with T threads per column
// has exactly two butterflies per thread (a mirror-block store can be used)
template <int L, int T, int MAXR>
constexpr bool mirror_last_stage() {
  constexpr RadixList rl = make_radices(L, MAXR);
  return rl.n > 1 && L / rl.r[rl.n - 1] == 2 * T;
}

constexpr int max_radix(int L, int maxr = 8) {
  RadixList rl = make_radices(L, maxr);
  int mx = 1;
  for (int i = 0; i < rl.n; ++i) mx = rl.r[i] > mx ? rl.r[i] : mx;
  return mx;
}

// threads per column: enough for one radix-max butterfly each in the
// largest-radix stages
template <int L, int MAXR = 8>
struct FftShape {
  static constexpr RadixList rl = make_radices(L, MAXR);
  static_assert(rl.n > 0, "FFT length must be 2^a 3^b 5^c");
  static constexpr int T = L / max_radix(L, MAXR);
  static constexpr int CS = L + L / 8 + 1;  // padded LDS column stride
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 3); }

// Staged global access of the first and last stages.  A plain Load is a
// callable load(i) -> complex and a plain Store a callable store(i, v).  A
// staged Load (with a member type `staged`) splits each element into
// fetch(i) -> raw_t, its global loads, and make(raw, i) -> complex, the
// arithmetic on them; a staged Store splits each output into prep(i) ->
// aux_t, the global loads it needs, and put(i, v, aux).  The stage issues
// the fetches (preps) of `group` elements of a butterfly before their makes
// (puts), so those loads are in flight together instead of one dependent
// load-compute chain per element; `group` bounds the registers they hold.
template <class T, class = void>
struct is_staged : std::false_type {};
template <class T>
struct is_staged<T, std::void_t<typename T::staged>> : std::true_type {};

// A mirror-block Store (with a member type `mirror_block`) takes the last
// stage's outputs in registers: that stage runs butterflies j and NB-1-j in
// thread j (it needs exactly two butterflies per thread), whose outputs are
// the elements j + r NS and (NS-1-j) + r NS, so element m and element L-1-m
// are in the same thread; it calls block(j, NB-1-j, v) once with
// v[0][r], v[1][r] the two butterflies' outputs (no LDS writes and no
// barrier).  mirror_last_stage<L, T, MAXR>() says whether a plan allows it.
template <class T, class = void>
struct is_mirror_block : std::false_type {};
template <class T>
struct is_mirror_block<T, std::void_t<typename T::mirror_block>> : std::true_type {};

// w[r] = w[1]^r for r = 2 .. R-1, each by one complex product of lower
// powers (at most three products deep, so within a few ulps of the
// correctly rounded table value).  One table load per butterfly instead of
// R - 1: the dependent global loads of the twiddles measured 10-20 % of the
// x and y passes at 512 points.
template <int R, class Re>
__device__ __forceinline__ void twiddle_powers(cx<Re> (&w)[R]) {
  if constexpr (R >= 3) w[2] = cmul(w[1], w[1]);
  if constexpr (R >= 4) w[3] = cmul(w[2], w[1]);
  if constexpr (R >= 5) w[4] = cmul(w[2], w[2]);
  if constexpr (R >= 6) w[5] = cmul(w[4], w[1]);
  if constexpr (R >= 7) w[6] = cmul(w[4], w[2]);
  if constexpr (R >= 8) w[7] = cmul(w[4], w[3]);
  static_assert(R <= 8, "radix above 8");
}

// One Stockham stage.  lds = this column's LDS region; tw = exp(-2 pi i q / L).
template <int L, int R, int NS, int S, int T, bool FIRST, bool LAST, class Re, class Load,
          class Store>
__device__ __forceinline__ void stockham_stage(cx<Re> *lds, int tid,
                                               const cx<Re> *__restrict__ tw,
                                               Load &load, Store &store) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
  constexpr bool MB = LAST && is_mirror_block<Store>::value;
  static_assert(!MB || NB == 2 * T, "a mirror-block store needs two butterflies per thread");
  cx<Re> v[PER][R];
  if constexpr (FIRST && is_staged<Load>::value) {
    // Load::group elements are fetched together (bounded by the registers
    // their raw values take)
    constexpr int G = (int)Load::group;
#pragma unroll
    for (int b = 0; b < PER; ++b) {
      const int j = tid + b * T;
      if ((NB % T == 0) || j < NB) {
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += G) {
          typename Load::raw_t raw[G];
#pragma unroll
          for (int g = 0; g < G; ++g)
            if (r0 + g < R) raw[g] = load.fetch(j + (r0 + g) * NB);
#pragma unroll
          for (int g = 0; g < G; ++g)
            if (r0 + g < R) v[b][r0 + g] = load.make(raw[g], j + (r0 + g) * NB);
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < PER; ++b) {
    const int j = MB ? (b == 0 ? tid : NB - 1 - tid) : tid + b * T;
    if ((NB % T == 0) || j < NB) {
      if constexpr (!(FIRST && is_staged<Load>::value)) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int idx = j + r * NB;
          if constexpr (FIRST)
            v[b][r] = load(idx);
          else
            v[b][r] = lds[lpad(idx)];
        }
      }
      if constexpr (NS > 1) {
        const int k = j % NS;
        cx<Re> w[R];
        w[1] = tw[k * (L / (NS * R))];
        twiddle_powers<R>(w);
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = (S < 0) ? cmul(v[b][r], w[r]) : cmulc(v[b][r], w[r]);
      }
      Dft<R, S>::run(v[b]);
    }
  }
  if constexpr (MB) {
    store.block(tid, NB - 1 - tid, v);
  } else {
    if constexpr (!FIRST) __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; ++b) {
      const int j = tid + b * T;
      if ((NB % T == 0) || j < NB) {
        const int k = j % NS;
        const int idxD = (j / NS) * NS * R + k;
        if constexpr (LAST && is_staged<Store>::value) {
          constexpr int G = (int)Store::group;
#pragma unroll
          for (int r0 = 0; r0 < R; r0 += G) {
            typename Store::aux_t aux[G];
#pragma unroll
            for (int g = 0; g < G; ++g)
              if (r0 + g < R) aux[g] = store.prep(idxD + (r0 + g) * NS);
#pragma unroll
            for (int g = 0; g < G; ++g)
              if (r0 + g < R) store.put(idxD + (r0 + g) * NS, v[b][r0 + g], aux[g]);
          }
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int idx = idxD + r * NS;
            if constexpr (LAST)
              store(idx, v[b][r]);
            else
              lds[lpad(idx)] = v[b][r];
          }
        }
      }
    }
    if constexpr (!LAST) __syncthreads();
  }
}

template <int L, int S, int T, int STAGE, int NS, bool FROM_GLOBAL, bool TO_GLOBAL, int MAXR,
          class Re, class Load, class Store>
__device__ __forceinline__ void stockham_run(cx<Re> *lds, int tid, const cx<Re> *__restrict__ tw,
                                             Load &load, Store &store) {
  constexpr RadixList rl = make_radices(L, MAXR);
  if constexpr (STAGE < rl.n) {
    constexpr int R = rl.r[STAGE];
    constexpr bool first = FROM_GLOBAL && STAGE == 0;
    constexpr bool last = TO_GLOBAL && STAGE == rl.n - 1;
    stockham_stage<L, R, NS, S, T, first, last>(lds, tid, tw, load, store);
    stockham_run<L, S, T, STAGE + 1, NS * R, FROM_GLOBAL, TO_GLOBAL, MAXR>(lds, tid, tw, load,
                                                                            store);
  }
}

// Full column FFT.  FROM_GLOBAL: stage 0 reads load(i); otherwise it reads the
// column from LDS (lds[lpad(i)]) and the caller must have synchronised.
// TO_GLOBAL: the last stage calls store(i, v); otherwise the result is left in
// LDS and the function ends with a barrier.  S = -1 forward, +1 inverse
// (unnormalised).
// TT threads work on the column (default FftShape<L>::T, one largest-radix
// butterfly each; fewer threads take several butterflies per stage).
// MAXR: the largest radix of the plan (8, or 4 for twice the threads).
template <int L, int S, bool FROM_GLOBAL, bool TO_GLOBAL, int TT = FftShape<L>::T, int MAXR = 8,
          class Re, class Load, class Store>
__device__ __forceinline__ void fft_column(cx<Re> *lds, int tid, const cx<Re> *__restrict__ tw,
                                           Load &load, Store &store) {
  stockham_run<L, S, TT, 0, 1, FROM_GLOBAL, TO_GLOBAL, MAXR>(lds, tid, tw, load, store);
}

}  // namespace rism
