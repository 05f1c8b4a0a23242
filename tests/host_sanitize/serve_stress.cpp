// Drives the library's host code built for a sanitizer against the CPU
// stand-in runtime (hip_stub.cpp): a CPD server with many admission
// workers and slots, requests submitted from several client threads (host
// arrays and binary files, FP32 and BF16 factors, 3- and 4-mode tensors,
// requests over the HBM budget, an injected allocation failure), waited
// for, fetched and released while others run; device-resident CPD solves on
// other threads at the same time.  Exit status 0 when every request ended
// as expected and the stand-in saw no misuse of streams, events or memory.
//
//   serve_stress [admit_threads] [slots] [requests]
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sptol_hip.h"

extern "C" int hipstub_misuse_count();

namespace {

struct Tensor {
    std::vector<uint32_t> nd;
    std::vector<std::vector<uint32_t>> inds;
    std::vector<float> vals;
    std::vector<sptIndexVector> iv;
    std::vector<uint32_t> so;
    sptSparseTensor t;
    Tensor(int nmodes, uint64_t nnz, uint32_t seed)
    {
        std::mt19937 rng(seed);
        nd.resize(nmodes);
        inds.resize(nmodes);
        for (int m = 0; m < nmodes; ++m) nd[m] = 20 + rng() % 300;
        for (int m = 0; m < nmodes; ++m) {
            inds[m].resize(nnz);
            for (auto &x : inds[m]) x = rng() % nd[m];
        }
        vals.resize(nnz);
        for (auto &v : vals) v = 0.01f + (float)(rng() % 1000) / 1000.f;
        iv.resize(nmodes);
        so.resize(nmodes);
        for (int m = 0; m < nmodes; ++m) {
            iv[m].len = iv[m].cap = nnz;
            iv[m].data = inds[m].data();
            so[m] = m;
        }
        t.nmodes = nmodes;
        t.sortorder = so.data();
        t.ndims = nd.data();
        t.nnz = nnz;
        t.inds = iv.data();
        t.values.len = t.values.cap = nnz;
        t.values.data = vals.data();
    }
};

std::atomic<int> g_bad{0};

void expect(bool ok, const char *what, long a = 0)
{
    if (!ok) {
        g_bad++;
        fprintf(stderr, "serve_stress: unexpected: %s (%ld)\n", what, a);
    }
}

void client(sptolServer srv, int id, int n, const std::string &dir)
{
    std::mt19937 rng(1000 + id);
    std::vector<uint64_t> ids;
    std::vector<Tensor *> keep;
    for (int k = 0; k < n; ++k) {
        const int nm = 3 + (int)(rng() % 2);
        const uint64_t nnz = 200 + rng() % 6000;
        auto *X = new Tensor(nm, nnz, (uint32_t)(id * 100000 + k));
        keep.push_back(X);
        const uint32_t R = (rng() % 2) ? 16 : 8;
        const int fdt = (rng() % 3 == 0) ? SPTOL_BF16 : SPTOL_F32;
        uint64_t rid = 0;
        int r;
        if (k % 7 == 3) {
            const std::string path = dir + "/c" + std::to_string(id) + "_" + std::to_string(k) + ".bin";
            expect(sptolSaveSparseTensorBinary(&X->t, path.c_str()) == 0, "save binary");
            r = sptolServerSubmitFile(srv, path.c_str(), R, 4 + rng() % 5, -1.0, k, fdt, &rid);
        } else {
            r = sptolServerSubmit(srv, &X->t, R, 4 + rng() % 5, (rng() % 2) ? 1e-4 : -1.0, k, fdt, &rid);
        }
        expect(r == 0, "submit", r);
        ids.push_back(rid);
        if (k % 5 == 0) {
            // a finished earlier request is fetched and released while others run
            const size_t j = rng() % ids.size();
            sptolRequestInfo info;
            if (ids[j] && sptolServerWait(srv, ids[j], 60.0, &info) == 0) {
                if (info.status == SPTOL_REQ_DONE) {
                    std::vector<float> lam(info.rank);
                    sptolServerFetch(srv, ids[j], nullptr, lam.data());
                }
                sptolServerRelease(srv, ids[j]);
                ids[j] = 0;
            }
        }
    }
    for (size_t j = 0; j < ids.size(); ++j) {
        if (!ids[j]) continue;
        sptolRequestInfo info;
        const int r = sptolServerWait(srv, ids[j], 120.0, &info);
        expect(r == 0, "wait", r);
        if (r == 0) expect(info.status == SPTOL_REQ_DONE || info.status == SPTOL_REQ_FAILED, "status", info.status);
        sptolServerRelease(srv, ids[j]);
    }
    for (auto *X : keep) delete X;
}

std::atomic<int> g_failed_solves{0};

void device_solves(int n)
{
    // plain device-resident CPD solves beside the server (shared block
    // cache); the one injected allocation failure may land here, so a solve
    // that fails is counted and cleaned up, and the test allows one
    for (int k = 0; k < n; ++k) {
        Tensor X(3, 3000, 777 + k);
        sptolDevCoo h = nullptr;
        sptolDevMatrix mats[3];
        std::memset(mats, 0, sizeof(mats));
        int r = sptolDevCooCreate(&h, &X.t);
        for (int m = 0; m < 3 && !r; ++m) {
            r = sptolDevMatrixCreate(&mats[m], X.nd[m], 8, SPTOL_F32);
            sptMatrix hm;
            sptNewMatrix(&hm, X.nd[m], 8);
            sptRandomizeMatrix(&hm, X.nd[m], 8);
            if (!r) r = sptolDevMatrixUpload(&mats[m], &hm);
            sptFreeMatrix(&hm);
        }
        float lam[8];
        sptolCpdStats st;
        std::memset(&st, 0, sizeof(st));
        st.exact_fit = k & 1;
        void *s = nullptr;
        if (!r) r = sptolHipStreamCreate(&s);
        if (!r) r = sptolCpdAlsDevice(h, 8, 3, -1.0, mats, lam, &st, s);
        if (s) sptolHipStreamDestroy(s);
        if (r) g_failed_solves++;
        for (int m = 0; m < 3; ++m) sptolDevMatrixFree(&mats[m]);
        if (h) sptolDevCooFree(h);
    }
}

}  // namespace

int main(int argc, char **argv)
{
    const int admit = argc > 1 ? std::atoi(argv[1]) : 8;
    const int slots = argc > 2 ? std::atoi(argv[2]) : 96;
    const int nreq = argc > 3 ? std::atoi(argv[3]) : 240;
    char dir[] = "/tmp/serve_stress_XXXXXX";
    if (!mkdtemp(dir)) return 2;
    sptolServerConfig cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.hbm_budget = (size_t)2 << 30;
    cfg.max_active = slots;
    cfg.batch_mttkrp = 1;
    cfg.exact_fit = 1;
    cfg.admit_threads = admit;
    sptolServer srv = nullptr;
    if (sptolServerCreate(&srv, &cfg) != 0) {
        fprintf(stderr, "server create failed\n");
        return 1;
    }
    // over the budget: rejected at admission
    {
        Tensor big(3, 1000, 5);
        uint64_t rid = 0;
        big.nd[0] = 2000000000u;
        big.t.ndims = big.nd.data();
        expect(sptolServerSubmit(srv, &big.t, 64, 3, -1.0, 1, SPTOL_F32, &rid) == 0, "submit big");
        sptolRequestInfo info;
        expect(sptolServerWait(srv, rid, 60.0, &info) == 0 && info.status == SPTOL_REQ_FAILED, "big rejected");
        sptolServerRelease(srv, rid);
    }
    // one allocation failure somewhere in the first requests
    sptolHipInjectMallocFailure(40);
    const int nclients = 3;
    std::vector<std::thread> th;
    for (int c = 0; c < nclients; ++c) th.emplace_back(client, srv, c, nreq / nclients, std::string(dir));
    std::thread dev(device_solves, 6);
    // kernel timing switched on while everything runs, its totals read
    // (and its finished records drained) over and over from another thread
    std::atomic<bool> stop_prof{false};
    std::thread prof([&] {
        sptolProfileEnable(1);
        while (!stop_prof.load()) {
            double ms = 0;
            uint64_t n = 0;
            sptolProfileQuery("mttkrp_seg", &ms, &n);
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        sptolProfileEnable(0);
        sptolProfileReset();
    });
    for (auto &t : th) t.join();
    dev.join();
    stop_prof = true;
    prof.join();
    expect(g_failed_solves.load() <= 1, "device solves failed", g_failed_solves.load());
    expect(sptolServerDrain(srv) == 0, "drain");
    fprintf(stderr, "serve_stress: %llu ticks\n", (unsigned long long)sptolServerTicks(srv));
    sptolServerDestroy(srv);
    sptolHipReleaseCache();
    const int misuse = hipstub_misuse_count();
    fprintf(stderr, "serve_stress: admit %d slots %d requests %d: %d unexpected, %d runtime misuses\n", admit, slots,
            nreq, g_bad.load(), misuse);
    std::string cmd = std::string("rm -rf ") + dir;
    (void)!system(cmd.c_str());
    return (g_bad.load() || misuse) ? 1 : 0;
}
