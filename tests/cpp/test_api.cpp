// test_api.cpp -- TAMM-style tests written against include/tamm/tamm.hpp (the C++ API over
// the MI355X engine).  They restate TAMM's own checks:
//   tests/tamm/Test_Ops.cpp:1400-1560     three/four-dimensional mult ops (closed forms)
//   tests/tamm/Test_Ops.cpp:1808-1880     MultOp with RHS reduction labels
//   tests/tamm/Test_Mult_Ops.cpp:161-206  4-D contraction, N=50 tile 20, norm == 2.625e8
// plus a spin-blocked ladder / ring contraction against a brute-force dense loop, the
// adjoint ops (tamm::adjoint) against the bilinearity identity <dV, V> = <dT, T> = <L, R>,
// Scheduler::capture replays against the eager level, and the kernels::block_multiply shim.  Exit status 0 = all checks passed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "tamm/tamm.hpp"

using namespace tamm;
using T = double;

static int g_fail = 0;
#define REQUIRE(cond)                                                        \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
      ++g_fail;                                                              \
    }                                                                        \
  } while (0)

template <typename TT>
static void check_value(Tensor<TT>& t, TT val) {
  std::vector<TT> buf;
  std::vector<Index> bid(t.num_modes(), 0);
  const auto& sp = t.tiled_index_spaces();
  while (true) {
    t.get(bid, buf);
    for (TT x : buf) REQUIRE(std::abs(x - val) < 1.0e-10);
    int d = (int)bid.size() - 1;
    for (; d >= 0; --d) {
      if (++bid[d] < (Index)sp[d].num_tiles()) break;
      bid[d] = 0;
    }
    if (d < 0) break;
  }
}

// Calls f(bid) for every nonzero block of t.
template <typename TT, typename F>
static void for_blocks(const Tensor<TT>& t, F f) {
  std::vector<Index> bid(t.num_modes(), 0);
  const auto& sp = t.tiled_index_spaces();
  while (true) {
    if (t.is_non_zero(bid)) f(bid);
    int d = (int)bid.size() - 1;
    for (; d >= 0; --d) {
      if (++bid[d] < (Index)sp[d].num_tiles()) break;
      bid[d] = 0;
    }
    if (d < 0) break;
  }
}

static double host_dot(Tensor<T>& x, Tensor<T>& y) {
  double s = 0;
  std::vector<T> bx, by;
  for_blocks(x, [&](const std::vector<Index>& bid) {
    x.get(bid, bx);
    y.get(bid, by);
    for (size_t e = 0; e < bx.size(); ++e) s += bx[e] * by[e];
  });
  return s;
}

static void fill_random(Tensor<T>& t, std::mt19937_64& gen, double scale) {
  std::normal_distribution<double> nd;
  for_blocks(t, [&](const std::vector<Index>& bid) {
    std::vector<T> buf(t.block_size(bid));
    for (auto& x : buf) x = scale * nd(gen);
    t.put(bid, buf);
  });
}

static void test_adjoint_identity(ExecutionContext& ec) {
  IndexSpace OS{range(0, 4), {}, {{Spin{1}, {range(0, 2)}}, {Spin{2}, {range(2, 4)}}}};
  IndexSpace VS{range(0, 8), {}, {{Spin{1}, {range(0, 4)}}, {Spin{2}, {range(4, 8)}}}};
  TiledIndexSpace O{OS, 2}, V{VS, 3};
  auto [a, b, c, d] = V.labels<4>("all");
  auto [i, j] = O.labels<2>("all");
  Tensor<T> Vt{{V, V, V, V}, {2, 2}}, Tt{{V, V, O, O}, {2, 2}}, Lt{{V, V, O, O}, {2, 2}};
  Tensor<T> R{{V, V, O, O}, {2, 2}}, dV{{V, V, V, V}, {2, 2}}, dT{{V, V, O, O}, {2, 2}};
  Scheduler{ec}.allocate(Vt, Tt, Lt, R, dV, dT).execute();
  std::mt19937_64 gen(11);
  fill_random(Vt, gen, 0.1);
  fill_random(Tt, gen, 0.01);
  fill_random(Lt, gen, 0.01);
  Scheduler sch{ec};
  sch(R() = 0.0)(dV() = 0.0)(dT() = 0.0)(R(a, b, i, j) += Vt(a, b, c, d) * Tt(c, d, i, j));
  for (auto& op : adjoint(Lt(a, b, i, j), Vt(a, b, c, d) * Tt(c, d, i, j), &dV, &dT)) sch(op);
  sch.execute(ExecutionHW::GPU);
  const double L = host_dot(Lt, R);
  const double gv = host_dot(dV, Vt), gt = host_dot(dT, Tt);
  std::printf("adjoint: L %.15e <dV,V> %.15e <dT,T> %.15e\n", L, gv, gt);
  REQUIRE(std::abs(L) > 0);
  REQUIRE(std::abs(gv - L) <= 1e-12 * std::abs(L));
  REQUIRE(std::abs(gt - L) <= 1e-12 * std::abs(L));
  Scheduler{ec}.deallocate(Vt, Tt, Lt, R, dV, dT).execute();
}

// Scheduler::capture: a level (SetOp + two MultOps) recorded into a hipGraph; replays after
// the output is overwritten reproduce the eager result bit for bit.
static void test_scheduler_capture(ExecutionContext& ec) {
  IndexSpace OS{range(0, 4), {}, {{Spin{1}, {range(0, 2)}}, {Spin{2}, {range(2, 4)}}}};
  IndexSpace VS{range(0, 8), {}, {{Spin{1}, {range(0, 4)}}, {Spin{2}, {range(4, 8)}}}};
  TiledIndexSpace O{OS, 2}, V{VS, 3};
  auto [a, b, c, d] = V.labels<4>("all");
  auto [i, j, k] = O.labels<3>("all");
  Tensor<T> Vt{{V, V, V, V}, {2, 2}}, Tt{{V, V, O, O}, {2, 2}}, Wt{{O, V, V, O}, {2, 2}}, R{{V, V, O, O}, {2, 2}};
  Scheduler{ec}.allocate(Vt, Tt, Wt, R).execute();
  std::mt19937_64 gen(5);
  fill_random(Vt, gen, 0.1);
  fill_random(Tt, gen, 0.01);
  fill_random(Wt, gen, 0.1);
  Scheduler sch{ec};
  sch(R() = 0.0)(R(a, b, i, j) += Vt(a, b, c, d) * Tt(c, d, i, j))(R(a, b, i, j) += Wt(k, b, c, j) * Tt(a, c, i, k));
  auto g = sch.capture();
  std::vector<std::vector<T>> eager;
  for_blocks(R, [&](const std::vector<Index>& bid) {
    std::vector<T> buf;
    R.get(bid, buf);
    eager.push_back(buf);
  });
  for (int rep = 0; rep < 3; ++rep) {
    for_blocks(R, [&](const std::vector<Index>& bid) {
      std::vector<T> junk(R.block_size(bid), 1.0e300);
      R.put(bid, junk);
    });
    g->launch();
    size_t n = 0;
    bool same = true;
    for_blocks(R, [&](const std::vector<Index>& bid) {
      std::vector<T> buf;
      R.get(bid, buf);
      same = same && buf == eager[n++];
    });
    REQUIRE(same);
  }
  g.reset();
  Scheduler{ec}.deallocate(Vt, Tt, Wt, R).execute();
}

static void test_mult_ops_3d_4d(ExecutionContext& ec) {
  IndexSpace IS{range(0, 10), {{"nr1", {range(0, 5)}}, {"nr2", {range(5, 10)}}}};
  TiledIndexSpace TIS{IS, 1};
  auto [i, j, k, l, m] = TIS.labels<5>("all");
  {
    Tensor<T> T1{TIS, TIS, TIS}, T2{TIS, TIS, TIS}, T3{TIS, TIS, TIS};
    Scheduler{ec}.allocate(T1, T2, T3)(T1() = 2)(T2() = 3)(T3() = 4)(T1(i, j, k) += 1.7 * T2(j, l, i) * T3(l, i, k))
        .execute(ExecutionHW::GPU);
    check_value(T1, T(2 + 1.7 * 3 * 4 * 10));
  }
  {
    Tensor<T> T1{TIS, TIS, TIS, TIS}, T2{TIS, TIS, TIS, TIS}, T3{TIS, TIS, TIS, TIS};
    Scheduler{ec}.allocate(T1, T2, T3)(T1() = 2)(T2() = 3)(T3() = 4)(T1(i, j, k, l) += 1.7 * T2(j, l, k, m) * T3(l, i, k, m))
        .execute(ExecutionHW::GPU);
    check_value(T1, T(2 + 1.7 * 3 * 4 * 10));
  }
  {
    Tensor<T> T1{TIS, TIS, TIS, TIS}, T2{TIS, TIS}, T3{TIS, TIS};
    Scheduler{ec}.allocate(T1, T2, T3)(T1() = 2)(T2() = 3)(T3() = 4)(T1(i, j, k, l) += 1.7 * T2(i, j) * T3(k, l))
        .execute(ExecutionHW::GPU);
    check_value(T1, T(2 + 1.7 * 3 * 4));
  }
}

static void test_rhs_reduction(ExecutionContext& ec) {
  IndexSpace IS{range(0, 4), {{"nr1", {range(0, 2)}}, {"nr2", {range(2, 4)}}}};
  TiledIndexSpace TIS{IS, 1};
  auto [i, j] = TIS.labels<2>("all");
  Tensor<T> T1{}, T2{TIS}, T3{TIS};
  Scheduler{ec}.allocate(T1, T2, T3)(T1() = 4)(T2() = 3)(T3() = 5)(T1() += -3.2 * T2(i) * T3(j))
      .execute(ExecutionHW::GPU);
  check_value(T1, 4.0 - 3.2 * 4 * 4 * 3 * 5);
}

// C(i,j) = A(i,k) * B(k,j) (TAMM's assignment form) with A and B spin-blocked ({1,1}) and C
// dense: C's mixed-spin blocks get no block pair and must read 0 after the assignment, the
// others 2 * 3 * 2 (k over the two orbitals of the block's spin), whatever C held before; run
// twice, so the second run must not accumulate.
static void test_assign_mult(ExecutionContext& ec) {
  IndexSpace OS{range(0, 4), {}, {{Spin{1}, {range(0, 2)}}, {Spin{2}, {range(2, 4)}}}};
  TiledIndexSpace O{OS, 2};
  auto [i, j, k] = O.labels<3>("all");
  Tensor<T> A{{O, O}, {1, 1}}, B{{O, O}, {1, 1}}, C{O, O};
  Scheduler{ec}.allocate(A, B, C)(A() = 2.0)(B() = 3.0)(C() = 5.0).execute();
  for (int rep = 0; rep < 2; ++rep) {
    Scheduler{ec}(C(i, j) = A(i, k) * B(k, j)).execute(ExecutionHW::GPU);
    for_blocks(C, [&](const std::vector<Index>& bid) {
      std::vector<T> buf;
      C.get(bid, buf);
      const T want = bid[0] == bid[1] ? T(12) : T(0);
      for (T x : buf) REQUIRE(x == want);
    });
  }
  Scheduler{ec}.deallocate(A, B, C).execute();
}

static void test_mult_ops_norm(ExecutionContext& ec) {
  TiledIndexSpace tis1{IndexSpace{range(50)}, 20};
  auto [i, j, k, l, m, o] = tis1.labels<6>("all");
  Tensor<T> A{i, j, m, o};
  Tensor<T> B{m, o, k, l};
  Tensor<T> C{i, j, k, l};
  Scheduler sch{ec};
  sch.allocate(A, B, C)(A() = 21.0)(B() = 2.0)(C() = 0.0).execute();
  sch(C(j, i, k, l) += A(i, j, m, o) * B(m, o, k, l)).execute(ExecutionHW::GPU, true);
  T n = norm(C);
  std::printf("Test_Mult_Ops norm %.17g (expected 2.625e8)\n", n);
  REQUIRE(n == 2.625e8);
}

// Spin-blocked ladder and ring on a small MO space against a dense brute-force loop.
static void test_spin_ladder_ring(ExecutionContext& ec) {
  const int o = 4, v = 8;
  IndexSpace OS{range(0, o), {}, {{Spin{1}, {range(0, o / 2)}}, {Spin{2}, {range(o / 2, o)}}}};
  IndexSpace VS{range(0, v), {}, {{Spin{1}, {range(0, v / 2)}}, {Spin{2}, {range(v / 2, v)}}}};
  TiledIndexSpace O{OS, 2}, V{VS, 3};
  auto [a, b, c, d] = V.labels<4>("all");
  auto [i, j, kk] = O.labels<3>("all");
  Tensor<T> Vt{{V, V, V, V}, {2, 2}}, Tt{{V, V, O, O}, {2, 2}}, Wt{{O, V, V, O}, {2, 2}}, R{{V, V, O, O}, {2, 2}};
  Scheduler{ec}.allocate(Vt, Tt, Wt, R).execute();
  // dense reference arrays filled through the block API
  std::mt19937_64 gen(7);
  std::normal_distribution<double> nd;
  auto spin = [](int p, int n) { return p < n / 2 ? 1 : 2; };
  std::vector<double> dV(v * v * v * v), dT(v * v * o * o), dW(o * v * v * o);
  for (auto& x : dV) x = 0.1 * nd(gen);
  for (auto& x : dT) x = 0.01 * nd(gen);
  for (auto& x : dW) x = 0.1 * nd(gen);
  auto put4 = [&](Tensor<T>& t, const std::vector<double>& dense, std::vector<int> n) {
    const auto& sp = t.tiled_index_spaces();
    std::vector<Index> bid(4, 0);
    std::vector<int> off(4);
    while (true) {
      if (t.is_non_zero(bid)) {
        auto dims = t.block_dims(bid);
        std::vector<double> buf(t.block_size(bid));
        for (int q = 0; q < 4; ++q) {
          off[q] = 0;
          for (Index s = 0; s < bid[q]; ++s) off[q] += (int)sp[q].tile_size(s);
        }
        size_t e = 0;
        for (size_t x0 = 0; x0 < dims[0]; ++x0)
          for (size_t x1 = 0; x1 < dims[1]; ++x1)
            for (size_t x2 = 0; x2 < dims[2]; ++x2)
              for (size_t x3 = 0; x3 < dims[3]; ++x3)
                buf[e++] = dense[(((off[0] + x0) * n[1] + off[1] + x1) * n[2] + off[2] + x2) * n[3] + off[3] + x3];
        t.put(bid, buf);
      }
      int dd = 3;
      for (; dd >= 0; --dd) {
        if (++bid[dd] < (Index)sp[dd].num_tiles()) break;
        bid[dd] = 0;
      }
      if (dd < 0) break;
    }
  };
  // zero the spin-forbidden elements of the dense arrays so both sides see the same data
  for (int p = 0; p < v; ++p)
    for (int q = 0; q < v; ++q)
      for (int r = 0; r < v; ++r)
        for (int s = 0; s < v; ++s)
          if (spin(p, v) + spin(q, v) != spin(r, v) + spin(s, v)) dV[((p * v + q) * v + r) * v + s] = 0;
  for (int p = 0; p < v; ++p)
    for (int q = 0; q < v; ++q)
      for (int r = 0; r < o; ++r)
        for (int s = 0; s < o; ++s)
          if (spin(p, v) + spin(q, v) != spin(r, o) + spin(s, o)) dT[((p * v + q) * o + r) * o + s] = 0;
  for (int p = 0; p < o; ++p)
    for (int q = 0; q < v; ++q)
      for (int r = 0; r < v; ++r)
        for (int s = 0; s < o; ++s)
          if (spin(p, o) + spin(q, v) != spin(r, v) + spin(s, o)) dW[((p * v + q) * v + r) * o + s] = 0;
  put4(Vt, dV, {v, v, v, v});
  put4(Tt, dT, {v, v, o, o});
  put4(Wt, dW, {o, v, v, o});
  Scheduler{ec}(R() = 0.0)(R(a, b, i, j) += Vt(a, b, c, d) * Tt(c, d, i, j))(R(a, b, i, j) += Wt(kk, b, c, j) * Tt(a, c, i, kk))
      .execute(ExecutionHW::GPU);
  std::vector<double> ref(v * v * o * o, 0.0);
  for (int pa = 0; pa < v; ++pa)
    for (int pb = 0; pb < v; ++pb)
      for (int pi = 0; pi < o; ++pi)
        for (int pj = 0; pj < o; ++pj) {
          double s = 0;
          for (int pc = 0; pc < v; ++pc)
            for (int pd = 0; pd < v; ++pd)
              s += dV[((pa * v + pb) * v + pc) * v + pd] * dT[((pc * v + pd) * o + pi) * o + pj];
          for (int pk = 0; pk < o; ++pk)
            for (int pc = 0; pc < v; ++pc)
              s += dW[((pk * v + pb) * v + pc) * o + pj] * dT[((pa * v + pc) * o + pi) * o + pk];
          ref[((pa * v + pb) * o + pi) * o + pj] = s;
        }
  std::vector<double> got(ref.size(), 0.0);
  const auto& sp = R.tiled_index_spaces();
  std::vector<Index> bid(4, 0);
  while (true) {
    std::vector<double> buf;
    R.get(bid, buf);
    auto dims = R.block_dims(bid);
    int off[4];
    for (int q = 0; q < 4; ++q) {
      off[q] = 0;
      for (Index s = 0; s < bid[q]; ++s) off[q] += (int)sp[q].tile_size(s);
    }
    size_t e = 0;
    for (size_t x0 = 0; x0 < dims[0]; ++x0)
      for (size_t x1 = 0; x1 < dims[1]; ++x1)
        for (size_t x2 = 0; x2 < dims[2]; ++x2)
          for (size_t x3 = 0; x3 < dims[3]; ++x3)
            got[(((off[0] + x0) * v + off[1] + x1) * o + off[2] + x2) * o + off[3] + x3] = buf[e++];
    int dd = 3;
    for (; dd >= 0; --dd) {
      if (++bid[dd] < (Index)sp[dd].num_tiles()) break;
      bid[dd] = 0;
    }
    if (dd < 0) break;
  }
  double num = 0, den = 0;
  for (size_t q = 0; q < ref.size(); ++q) {
    num += (got[q] - ref[q]) * (got[q] - ref[q]);
    den += ref[q] * ref[q];
  }
  double rel = std::sqrt(num / den);
  std::printf("spin ladder+ring vs dense loop: rel %.3e\n", rel);
  REQUIRE(rel < 1e-12);
}

// TAMM's strong size type (StrongNum<OffsetSpace, uint64_t>, types.hpp:73-89) and GPU stream
// handle (std::pair<hipStream_t, rocblas_handle>, gpu_streams.hpp:31), as direct callers pass
// them (Test_CCSD.cpp:425-441).
struct StrongSize {
  uint64_t v;
  uint64_t value() const { return v; }
};
struct RocblasHandleStandIn;
using TestGpuStream = std::pair<void*, RocblasHandleStandIn*>;

static void test_block_multiply_shim() {
  // C(j,i) = beta * C(j,i) + alpha * sum_k A(i,k) * B(k,j) for one block, called with TAMM's
  // argument types; beta = 0 overwrites (Test_CCSD passes cscale = 0), 1 accumulates.
  const size_t ni = 37, nj = 19, nk = 23;
  std::vector<double> A(ni * nk), B(nk * nj), C0(ni * nj);
  for (size_t q = 0; q < A.size(); ++q) A[q] = std::sin(0.1 * q);
  for (size_t q = 0; q < B.size(); ++q) B[q] = std::cos(0.07 * q);
  for (size_t q = 0; q < C0.size(); ++q) C0[q] = 0.5 + std::sin(0.3 * q);
  void *da, *db, *dc;
  detail::check(tmx_malloc(A.size() * 8, &da));
  detail::check(tmx_malloc(B.size() * 8, &db));
  detail::check(tmx_malloc(ni * nj * 8, &dc));
  double* th_a = static_cast<double*>(da);
  double* th_b = static_cast<double*>(db);
  double* cdev = static_cast<double*>(dc);
  double* ctmp = nullptr;
  void* engine_stream = nullptr;
  detail::check(tmx_get_stream(&engine_stream));
  TestGpuStream thandle{engine_stream, nullptr};
  std::vector<StrongSize> adims_sz{{ni}, {nk}}, bdims_sz{{nk}, {nj}}, cdims_sz{{nj}, {ni}};
  IntLabelVec alab{-1, -3}, blab{-3, -2}, clab{-2, -1};
  std::vector<double> cbuf(ni * nj, 0.0);
  for (double cscale : {0.0, 1.0, 0.25}) {
    detail::check(tmx_memcpy(cdev, C0.data(), C0.size() * 8, TMX_H2D, nullptr));
    kernels::block_multiply<T, T, T, T>(th_a, th_b, thandle, 1.5, A.data(), adims_sz, alab, B.data(), bdims_sz,
                                         blab, cscale, cbuf.data(), cdims_sz, clab, ExecutionHW::GPU, false, cdev,
                                         ctmp);
    std::vector<double> got(ni * nj);
    detail::check(tmx_memcpy(got.data(), cdev, got.size() * 8, TMX_D2H, nullptr));
    double err = 0;
    for (size_t j = 0; j < nj; ++j)
      for (size_t i = 0; i < ni; ++i) {
        double s = 0;
        for (size_t k = 0; k < nk; ++k) s += A[i * nk + k] * B[k * nj + j];
        const double ref = cscale * C0[j * ni + i] + 1.5 * s;
        err = std::max(err, std::abs(got[j * ni + i] - ref));
      }
    std::printf("block_multiply shim (gpuStream_t, Size dims) beta %.2f max err %.3e\n", cscale, err);
    REQUIRE(err < 1e-12);
  }
  tmx_free(da);
  tmx_free(db);
  tmx_free(dc);
}

// Per-op ExecutionHW (scheduler.hpp:40-51, :249-250) and the profile table
// (execution_context.hpp:358-365): a CPU op is refused before anything runs; levels follow
// the read/write dependences.
static void test_scheduler_hw_and_profile(ExecutionContext& ec) {
  TiledIndexSpace tis{IndexSpace{range(12)}, 4};
  auto [i, j, k] = tis.labels<3>("all");
  Tensor<T> A{tis, tis}, B{tis, tis}, C{tis, tis}, D{tis, tis};
  Scheduler{ec}.allocate(A, B, C, D)(A() = 1.0)(B() = 2.0)(C() = 0.0)(D() = 0.0).execute();
  bool threw = false;
  try {
    Scheduler{ec}(C(i, j) += A(i, k) * B(k, j), "cpu_op", ExecutionHW::CPU).execute(ExecutionHW::GPU);
  } catch (const std::runtime_error&) {
    threw = true;
  }
  REQUIRE(threw);
  check_value(C, 0.0);  // nothing ran
  ec.get_profile_data().str("");
  Scheduler sch{ec};
  sch(C(i, j) += A(i, k) * B(k, j), "c1", ExecutionHW::GPU)(C(i, j) += 0.5 * A(i, k) * B(k, j), "c2")(
      D(i, j) += C(i, k) * A(k, j), "d1");
  auto order = sch.levelize_and_order();
  REQUIRE(order.size() == 3 && order[0].first == 0 && order[1].first == 0 && order[2].first == 1);
  sch.execute(ExecutionHW::DEFAULT, true);
  check_value(C, T(1.5 * 2 * 12));
  check_value(D, T(1.5 * 2 * 12 * 12));
  const std::string table = ec.get_profile_data().str();
  std::printf("profile header: %s\n%s", ec.get_profile_header().c_str(), table.c_str());
  REQUIRE(table.find("0;0;c1;") == 0);
  REQUIRE(table.find("\n1;0;c2;") != std::string::npos);
  REQUIRE(table.find("\n2;1;d1;") != std::string::npos);
  REQUIRE(table.find(";SUM;") != std::string::npos);
  Scheduler{ec}.deallocate(A, B, C, D).execute();
}

int main(int argc, char** argv) {
  tamm::initialize(argc, argv, 0.25);
  {
    ExecutionContext ec;
    test_mult_ops_3d_4d(ec);
    test_rhs_reduction(ec);
    test_assign_mult(ec);
    test_mult_ops_norm(ec);
    test_spin_ladder_ring(ec);
    test_adjoint_identity(ec);
    test_scheduler_capture(ec);
    test_block_multiply_shim();
    test_scheduler_hw_and_profile(ec);
    bool threw = false;
    try {
      TiledIndexSpace tis{IndexSpace{range(10)}, 5};
      Tensor<T> X{tis};
      Scheduler{ec}.allocate(X)(X() = 1.0).execute(ExecutionHW::CPU);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    REQUIRE(threw);
  }
  tamm::finalize();
  std::printf(g_fail ? "test_api: %d FAILED\n" : "test_api: all passed\n", g_fail);
  return g_fail ? 1 : 0;
}
