// A dawgs-style C++ consumer of libp::ogs::ogs_t (the API of the reference's
// include/ogs.hpp:216-344, served by libogs_hip.so): the calls a PerformanceTest
// / CorrectnessTest makes (dawgsMain.cpp:146-416) on one rank, with the
// results written out for the GPU test to check against the oracle.
//
//   ogs_consumer nx ny nz p seed outdir
//
// Writes (raw little-endian): ids.i64 (the numbering), q.f64 (input),
// gs.f64 (Unsigned GatherScatter Add Sym), gather.f64 (Signed-unique Gather
// Add Trans), scatter.f64 (its Scatter NoTrans into zeros), gs_max.i32
// (GatherScatter Max of the local index as int), and prints "undef caught"
// when Gather on a non-unique ogs throws GATHER_UNDEF.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "ogs.hpp"

using namespace libp;

// Structured numbering of an nx x ny x nz box of degree-p hexes at one rank
// (the rule of dawgsMain.cpp:124-134): 1-based global ids, local index
// element-major, nodes i fastest.
static memory<hlong> make_ids(int nx, int ny, int nz, int p) {
  const int Nq = p + 1;
  const hlong sx = (hlong)p * nx + 1, sy = (hlong)p * ny + 1;
  memory<hlong> ids((size_t)nx * ny * nz * Nq * Nq * Nq);
  size_t n = 0;
  for (int K = 0; K < nz; ++K)
    for (int J = 0; J < ny; ++J)
      for (int I = 0; I < nx; ++I)
        for (int k = 0; k < Nq; ++k)
          for (int j = 0; j < Nq; ++j)
            for (int i = 0; i < Nq; ++i)
              ids[n++] = (hlong)(I * p + i) + (hlong)(J * p + j) * sx + (hlong)(K * p + k) * sx * sy + 1;
  return ids;
}

template <typename T>
static void dump(const std::string &path, const T *p, size_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(f);
}

template <typename T>
static std::vector<T> fetch(const deviceMemory<T> &d) {
  std::vector<T> h(d.length());
  LIBP_HIP_CHECK(hipMemcpy(h.data(), d.ptr(), h.size() * sizeof(T), hipMemcpyDeviceToHost));
  return h;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s nx ny nz p seed outdir\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[1]), ny = std::atoi(argv[2]), nz = std::atoi(argv[3]), p = std::atoi(argv[4]);
  const unsigned seed = (unsigned)std::atoi(argv[5]);
  const std::string out = argv[6];
  try {
    platform_t platform;
    memory<hlong> ids = make_ids(nx, ny, nz, p);
    const dlong N = (dlong)ids.length();
    dump(out + "/ids.i64", ids.ptr(), ids.length());

    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    memory<double> q((size_t)N);
    for (dlong n = 0; n < N; ++n) q[n] = U(gen);
    dump(out + "/q.f64", q.ptr(), q.length());

    // GatherScatter on the Unsigned ogs (dawgsMain.cpp:211-218)
    ogs::ogs_t ogs;
    ogs.Setup(N, ids, comm_t(), ogs::Unsigned, ogs::Auto, false, false, platform);
    deviceMemory<double> o_q = platform.malloc<double>((size_t)N, q.ptr());
    ogs.GatherScatter(o_q, 1, ogs::Add, ogs::Sym);
    platform.finish();
    dump(out + "/gs.f64", fetch(o_q).data(), (size_t)N);

    // Max over int values through the same handle
    memory<int32_t> idx((size_t)N);
    for (dlong n = 0; n < N; ++n) idx[n] = n;
    deviceMemory<int32_t> o_i = platform.malloc<int32_t>((size_t)N, idx.ptr());
    ogs.GatherScatter(o_i, 1, ogs::Max, ogs::Sym);
    platform.finish();
    dump(out + "/gs_max.i32", fetch(o_i).data(), (size_t)N);

    // Gather on a non-unique ogs is undefined (ogsSetup.cpp:861-864)
    try {
      deviceMemory<double> o_g = platform.malloc<double>((size_t)ogs.Ngather);
      ogs.Gather(o_g, o_q, 1, ogs::Add, ogs::Trans);
      std::printf("undef not caught\n");
    } catch (const libp::exception &e) {
      if (e.code() == OGS_E_GATHER_UNDEF) std::printf("undef caught\n");
      else std::printf("undef wrong code %d\n", e.code());
    }

    // Gather Trans / Scatter NoTrans on the Signed-unique sogs (dawgsMain.cpp:231-300)
    memory<hlong> sids = make_ids(nx, ny, nz, p);
    ogs::ogs_t sogs;
    sogs.Setup(N, sids, comm_t(), ogs::Signed, ogs::Auto, true, false, platform);
    deviceMemory<double> o_x = platform.malloc<double>((size_t)N, q.ptr());
    deviceMemory<double> o_gq = platform.malloc<double>((size_t)sogs.NlocalT);
    sogs.Gather(o_gq, o_x, 1, ogs::Add, ogs::Trans);
    std::vector<double> zeros((size_t)N, 0.0);
    deviceMemory<double> o_s = platform.malloc<double>((size_t)N, zeros.data());
    sogs.Scatter(o_s, o_gq, 1, ogs::NoTrans);
    platform.finish();
    dump(out + "/gather.f64", fetch(o_gq).data(), (size_t)sogs.NlocalT);
    dump(out + "/scatter.f64", fetch(o_s).data(), (size_t)N);
    std::printf("ok N=%d Ngather=%d NgatherGlobal=%lld\n", (int)N, (int)ogs.Ngather, (long long)ogs.NgatherGlobal);
  } catch (const libp::exception &e) {
    std::fprintf(stderr, "%s (code %d)\n", e.what(), e.code());
    return 1;
  }
  return 0;
}
