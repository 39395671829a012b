// Instantiates the column- and row-pass shapes whose transform length lies in
// [NS2D_TU_LO, NS2D_TU_HI] (set on the compiler command line; the build
// compiles this file once per range so that the ranges compile in parallel).
// Shapes are appended in the order of NS2D_KPASS_VARIANTS /
// NS2D_RPASS_VARIANTS, which fixes their per-length indices.
#include "ns2d_kernels.hpp"

#if !defined(NS2D_TU_LO) || !defined(NS2D_TU_HI)
#error "NS2D_TU_LO and NS2D_TU_HI must be defined"
#endif

namespace ns2d {

namespace {

template <bool ON, int L, int C, int T, int KS, int PS, int OFF, int PK, int W, int PF>
struct KAdd {
  static void add(std::vector<KVariant>&) {}
};
template <int L, int C, int T, int KS, int PS, int OFF, int PK, int W, int PF>
struct KAdd<true, L, C, T, KS, PS, OFF, PK, W, PF> {
  static void add(std::vector<KVariant>& v) {
    v.push_back({L, C, T, KS, PS, OFF, PK, W, PF, kpass_kernel<L, C, T, KS, PS, OFF, PK, W, PF>,
                 kpass_batched<L, C, T, KS, PS, OFF, PK, W, PF>});
  }
};

template <bool ON, int NY, int TLR, int PAIRS, int W, int PFR>
struct RAdd {
  static void add(std::vector<RVariant>&) {}
};
template <int NY, int TLR, int PAIRS, int W, int PFR>
struct RAdd<true, NY, TLR, PAIRS, W, PFR> {
  static void add(std::vector<RVariant>& v) {
    v.push_back({NY, TLR, PAIRS, W, PFR, rpass_kernel<NY, TLR, PAIRS, W, PFR>, rpass_batched<NY, TLR, PAIRS, W, PFR>,
                 rpass_cfl_kernel<NY, TLR, PAIRS, W, PFR>});
  }
};

constexpr bool in_tu(int n) { return n >= NS2D_TU_LO && n <= NS2D_TU_HI; }

}  // namespace

#define NS2D_CAT4_(a, b, c, d) a##b##_##d
#define NS2D_CAT4(a, b, c, d) NS2D_CAT4_(a, b, c, d)

void NS2D_CAT4(kv_add_, NS2D_TU_LO, _, NS2D_TU_HI)(std::vector<KVariant>& v) {
#define NS2D_KV_ADD(L_, C_, T_, KS_, PS_, OFF_, PK_, W_, PF_) \
  KAdd<in_tu(L_), L_, C_, T_, KS_, PS_, OFF_, PK_, W_, PF_>::add(v);
  NS2D_KPASS_VARIANTS(NS2D_KV_ADD)
#undef NS2D_KV_ADD
}

void NS2D_CAT4(rv_add_, NS2D_TU_LO, _, NS2D_TU_HI)(std::vector<RVariant>& v) {
#define NS2D_RV_ADD(NY_, TLR_, PAIRS_, W_, PFR_) RAdd<in_tu(NY_), NY_, TLR_, PAIRS_, W_, PFR_>::add(v);
  NS2D_RPASS_VARIANTS(NS2D_RV_ADD)
#undef NS2D_RV_ADD
}

}  // namespace ns2d
