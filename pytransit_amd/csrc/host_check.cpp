// Host-side checks of the C-ABI of librr_hip.so, for a build of the library's
// sources with AddressSanitizer and UndefinedBehaviorSanitizer on the host code
// (Makefile.sanitize).  Every call below returns before any HIP runtime call or
// kernel launch, so the program runs without a GPU: it drives the argument
// validation, the workspace arithmetic, the z grid and the host-side planning of
// request batches (with stand-in device addresses that the host never reads).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/rr_hip.h"

static int failures = 0;

#define CHECK(cond)                                                                        \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            std::fprintf(stderr, "FAILED %s:%d: %s (%s)\n", __FILE__, __LINE__, #cond,      \
                         rr_last_error());                                                 \
            ++failures;                                                                    \
        }                                                                                  \
    } while (0)

static void* fake(uintptr_t tag) { return (void*)(0x100000000ull + (tag << 20)); }

int main() {
    // z grid: create_z_grid's closed-form values and its argument checks
    std::vector<double> ze(40), zm(40), mu(40);
    CHECK(rr_z_grid(0.7, 20, 20, ze.data(), zm.data(), mu.data()) == RR_OK);
    CHECK(std::fabs(ze[0] - 0.035) < 1e-15 && std::fabs(ze[19] - 0.7) < 1e-15 && ze[39] == 1.0 && zm[0] == 0.0);
    CHECK(rr_z_grid(1.5, 2, 2, ze.data(), zm.data(), mu.data()) == RR_EINVAL);
    CHECK(rr_z_grid(0.7, 20, 20, nullptr, zm.data(), mu.data()) == RR_EINVAL);

    // workspace sizes
    CHECK(rr_workspace_bytes(4096, 20000, 1, 100, 1) >= 4096u * (192 + 8 + 100 * 12));
    CHECK(rr_workspace_bytes(-1, 10, 1, 100, 1) == 0);
    CHECK(ma_workspace_bytes(-3) == 0 && ec_workspace_bytes(-3) == 0);
    CHECK(ec_workspace_bytes(10) >= 10u * 128);

    // null and malformed descriptors
    CHECK(rr_evaluate(RR_MIXED, nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr) == RR_EINVAL);
    CHECK(rr_ld_profiles(42, nullptr, 1, 1, 2, nullptr, 4, nullptr, nullptr, nullptr) == RR_EINVAL);
    CHECK(eclipse_evaluate(nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr) == RR_EINVAL);
    CHECK(op_evaluate(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1, nullptr, 0, nullptr, nullptr) ==
          RR_ERANGE);
    CHECK(rr_evaluate_batch(RR_MIXED, nullptr, 17, nullptr, nullptr) == RR_EINVAL);
    CHECK(rr_evaluate_batch(RR_MIXED, nullptr, 0, nullptr, nullptr) == RR_OK);

    // a batch of 16 requests planned on the host with stand-in device addresses
    rr_table tab;
    std::memset(&tab, 0, sizeof(tab));
    tab.nk = 256; tab.ng = 100; tab.nz = 40;
    tab.kmin = 0.005; tab.kmax = 0.5; tab.dk = (tab.kmax - tab.kmin) / tab.nk; tab.dg = 1.0 / 99;
    tab.weights = (const double*)fake(1); tab.weights32 = (const float*)fake(2);
    tab.ze = (const double*)fake(3); tab.mu = (const double*)fake(4);
    std::vector<rr_request> reqs(16);
    for (int r = 0; r < 16; ++r) {
        rr_request& q = reqs[r];
        std::memset(&q, 0, sizeof(q));
        const int npb = 1 + r % 5, npv = 8 + 37 * r, npt = 1000 + 997 * r;
        q.lc.npt = npt; q.lc.nlc = npb; q.lc.npb = npb; q.lc.nep = 1;
        q.lc.sorted = RR_SORTED_RUNS | RR_SORTED_TIMES | RR_ONE_WINDOW;
        q.lc.uniform_nsamples = r % 3 == 0 ? 10 : 1;
        q.lc.times = (const double*)fake(10 + r); q.lc.lcids = (const int32_t*)fake(30 + r);
        q.lc.pbids = (const int32_t*)fake(50 + r); q.lc.epids = (const int32_t*)fake(70 + r);
        q.lc.nsamples = (const int32_t*)fake(90 + r); q.lc.exptimes = (const double*)fake(110 + r);
        q.pv.npv = npv; q.pv.kcols = 1; q.pv.t0cols = 1;
        q.pv.k = q.pv.t0 = q.pv.p = q.pv.a = q.pv.inc = q.pv.e = q.pv.w = (const double*)fake(130 + r);
        q.ld.law = r % 2 ? RR_LD_QUADRATIC : RR_LD_POWER_2; q.ld.nldc = 2; q.ld.ldc = (const double*)fake(150 + r);
        q.workspace = fake(170 + r);
        q.workspace_bytes = rr_workspace_bytes(npv, npt, npb, tab.ng, npb);
        q.flux = fake(190 + r);
    }
    std::vector<unsigned char> desc(rr_batch_desc_bytes());
    rr_batch_shape sh;
    CHECK(rr_batch_plan(RR_MIXED, &tab, 16, reqs.data(), desc.data(), &sh) == RR_OK);
    CHECK(sh.setup_blocks > 0 && sh.setup_multi == 1 && sh.flux_blocks[1] > 0 && sh.flux_blocks[0] == 0);
    CHECK(sh.setup_lds > 0 && sh.flux_lds[1] > 0);
    reqs[7].workspace_bytes = 16;   // a workspace too small: the whole batch is refused
    CHECK(rr_batch_plan(RR_MIXED, &tab, 16, reqs.data(), desc.data(), &sh) == RR_EINVAL);
    CHECK(std::strstr(rr_last_error(), "[7]") != nullptr);
    CHECK(rr_batch_launch(RR_MIXED, desc.data(), nullptr, nullptr) == RR_EINVAL);

    // the allocation-free entry point refuses problems above its scratch array
    rr_lcdata lc;
    std::memset(&lc, 0, sizeof(lc));
    lc.npt = 1000; lc.nlc = 1; lc.npb = 1; lc.nep = 1;
    rr_params pv;
    std::memset(&pv, 0, sizeof(pv));
    pv.npv = 50000; pv.kcols = 1; pv.t0cols = 1;
    rr_ld ld;
    std::memset(&ld, 0, sizeof(ld));
    ld.law = RR_LD_QUADRATIC; ld.nldc = 2;
    CHECK(rr_evaluate_noalloc(RR_FP64, &tab, &lc, &pv, &ld, nullptr, nullptr) == RR_ERANGE);

    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("host checks passed\n");
    return 0;
}
