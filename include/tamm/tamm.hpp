// tamm/tamm.hpp -- TAMM's labeled-tensor C++ API over the MI355X engine (header only).
//
// Source-compatible subset of TAMM's user API for the GPU contraction path:
//   IndexSpace{range, named subspaces, spin ranges}        src/tamm/index_space.hpp:64-82
//   TiledIndexSpace{is, tile} / {is, tiles}, tis("occ"),   src/tamm/tiled_index_space.hpp:31,49
//     tis.labels<N>("name")                                 :175
//   Tensor<T>{tis...}, Tensor<T>{{tis...}, {nup, nlo}}      src/tamm/tensor.hpp:115-134
//   C(a,b) = / += alpha * A(..) * B(..), C() = v, C(..) += alpha * A(..)
//                                                           src/tamm/labeled_tensor.hpp:62-75,274-290
//   Scheduler{ec}(op)...execute(ExecutionHW::GPU)          src/tamm/scheduler.hpp:40-51,160
//   tamm::norm(T)                                           src/tamm/tamm_utils.hpp:1197-1254
//   kernels::block_multiply (GPU signature)                 src/tamm/kernels/multiply.hpp:212-220
// Every operation runs on the GPU through include/tamm_mi355x.h; ExecutionHW::CPU is
// rejected (the CPU path stays TAMM's own).  Errors from the engine surface as
// std::runtime_error, as TAMM's HIP_CHECK does (src/tamm/gpu_streams.hpp:38-47).
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <initializer_list>
#include <type_traits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "tamm_mi355x.h"

namespace tamm {

using Tile = uint32_t;
using Index = uint32_t;
using IndexVector = std::vector<Index>;
using SizeVec = std::vector<size_t>;
using IntLabelVec = std::vector<int>;

enum class ExecutionHW { CPU, GPU, DEFAULT };

namespace detail {
inline void check(int rc) {
  if (rc != TMX_OK) throw std::runtime_error(std::string("tamm-mi355x: ") + tmx_last_error());
}
inline int next_label_id() {
  static int id = 0;
  return --id;  // negative ints, as MultOp::fillin_int_labels assigns (multop.hpp:955-970)
}
}  // namespace detail

inline void initialize(int /*argc*/ = 0, char** /*argv*/ = nullptr, double pool_fraction = 0.0) {
  detail::check(tmx_init(0, pool_fraction));
}
inline void finalize() { detail::check(tmx_finalize()); }

struct Range {
  int64_t lo, hi;
};
inline Range range(int64_t lo, int64_t hi) { return {lo, hi}; }
inline Range range(int64_t n) { return {0, n}; }

struct Spin {
  int v;
  explicit Spin(int s) : v(s) {}
  bool operator<(const Spin& o) const { return v < o.v; }
};
struct Spatial {
  int v;
  explicit Spatial(int s) : v(s) {}
  bool operator<(const Spatial& o) const { return v < o.v; }
};

class IndexSpace {
 public:
  using NamedRanges = std::vector<std::pair<std::string, std::vector<Range>>>;
  IndexSpace(Range r, NamedRanges named = {}, std::vector<std::pair<Spin, std::vector<Range>>> spin = {},
             std::vector<std::pair<Spatial, std::vector<Range>>> spatial = {})
      : range_(r), named_(std::move(named)) {
    for (auto& s : spin)
      for (auto& rr : s.second) spin_.push_back({rr.lo, rr.hi, s.first.v});
    for (auto& s : spatial)
      for (auto& rr : s.second) spatial_.push_back({rr.lo, rr.hi, s.first.v});
  }
  int64_t num_indices() const { return range_.hi - range_.lo; }

 private:
  friend class TiledIndexSpace;
  struct Attr { int64_t lo, hi; int v; };
  Range range_;
  NamedRanges named_;
  std::vector<Attr> spin_, spatial_;
};

class TiledIndexLabel;

// A tiled space, or a named subspace of one (a tile range of its root).
class TiledIndexSpace {
 public:
  TiledIndexSpace() = default;
  TiledIndexSpace(const IndexSpace& is, Tile tile) { build(is, 0, {tile}); }
  TiledIndexSpace(const IndexSpace& is, const std::vector<Tile>& tiles) { build(is, 1, tiles); }

  TiledIndexSpace operator()(const std::string& name = "all") const {
    if (name == "all") return *this;
    auto it = root_->sub.find(name);
    if (it == root_->sub.end()) throw std::runtime_error("no subspace named " + name);
    TiledIndexSpace s = *this;
    s.t0_ = it->second.first;
    s.nt_ = it->second.second;
    return s;
  }
  template <size_t N>
  auto labels(const std::string& name = "all") const;
  TiledIndexLabel label(const std::string& name = "all") const;

  int64_t num_tiles() const { return nt_; }
  int64_t tile_offset() const { return t0_; }
  size_t tile_size(Index t) const { return root_->toff[t0_ + t + 1] - root_->toff[t0_ + t]; }
  bool same_root(const TiledIndexSpace& o) const { return root_ == o.root_; }
  // Engine space holding exactly this space's tiles (created on demand, cached).
  tmx_space handle() const {
    auto key = std::make_pair(t0_, nt_);
    auto it = root_->views.find(key);
    if (it != root_->views.end()) return it->second;
    std::vector<int64_t> tiles;
    for (int64_t t = t0_; t < t0_ + nt_; ++t) tiles.push_back(root_->toff[t + 1] - root_->toff[t]);
    int64_t lo = root_->toff[t0_], n = root_->toff[t0_ + nt_] - lo;
    std::vector<int64_t> slo, shi, ilo, ihi;
    std::vector<int> sv, iv;
    for (int64_t t = t0_; t < t0_ + nt_; ++t) {
      slo.push_back(root_->toff[t] - lo);
      shi.push_back(root_->toff[t + 1] - lo);
      sv.push_back(root_->spin[t]);
      if (root_->irrep[t] >= 0) {
        ilo.push_back(root_->toff[t] - lo);
        ihi.push_back(root_->toff[t + 1] - lo);
        iv.push_back(root_->irrep[t]);
      }
    }
    bool has_spin = root_->has_spin;
    tmx_space s = nullptr;
    detail::check(tmx_space_create(n, 1, tiles.data(), (int64_t)tiles.size(), has_spin ? (int)slo.size() : 0,
                                   slo.data(), shi.data(), sv.data(), (int)ilo.size(), ilo.data(),
                                   ihi.data(), iv.data(), 0, nullptr, nullptr, &s));
    root_->views[key] = s;
    return s;
  }

 private:
  struct Root {
    std::vector<int64_t> toff;
    std::vector<int> spin, irrep;
    bool has_spin = false;
    std::map<std::string, std::pair<int64_t, int64_t>> sub;  // name -> (first tile, tiles)
    std::map<std::pair<int64_t, int64_t>, tmx_space> views;
    ~Root() {
      for (auto& kv : views) tmx_space_destroy(kv.second);
    }
  };
  void build(const IndexSpace& is, int use_vector, const std::vector<Tile>& tiles) {
    std::vector<int64_t> t(tiles.begin(), tiles.end()), slo, shi, ilo, ihi, nlo, nhi;
    std::vector<int> sv, iv;
    for (auto& a : is.spin_) { slo.push_back(a.lo); shi.push_back(a.hi); sv.push_back(a.v); }
    for (auto& a : is.spatial_) { ilo.push_back(a.lo); ihi.push_back(a.hi); iv.push_back(a.v); }
    for (auto& nr : is.named_)
      for (auto& r : nr.second) { nlo.push_back(r.lo); nhi.push_back(r.hi); }
    tmx_space s = nullptr;
    detail::check(tmx_space_create(is.num_indices(), use_vector, t.data(), (int64_t)t.size(), (int)slo.size(),
                                   slo.data(), shi.data(), sv.data(), (int)ilo.size(), ilo.data(), ihi.data(),
                                   iv.data(), (int)nlo.size(), nlo.data(), nhi.data(), &s));
    root_ = std::make_shared<Root>();
    int64_t nt = 0;
    detail::check(tmx_space_ntiles(s, &nt));
    root_->toff.resize(nt + 1);
    root_->spin.resize(nt);
    root_->irrep.resize(nt);
    detail::check(tmx_space_tiles(s, root_->toff.data(), root_->spin.data(), root_->irrep.data()));
    root_->has_spin = !is.spin_.empty();
    root_->views[{0, nt}] = s;
    t0_ = 0;
    nt_ = nt;
    for (auto& nr : is.named_) {
      // a named subspace is the tile range covering its index ranges (tiles split there)
      int64_t lo = nr.second.front().lo, hi = nr.second.back().hi;
      int64_t a = 0, b = 0;
      while (a < nt && root_->toff[a] < lo) ++a;
      b = a;
      while (b < nt && root_->toff[b] < hi) ++b;
      root_->sub[nr.first] = {a, b - a};
    }
  }
  std::shared_ptr<Root> root_;
  int64_t t0_ = 0, nt_ = 0;
};

class TiledIndexLabel {
 public:
  TiledIndexLabel() = default;
  TiledIndexLabel(const TiledIndexSpace& tis, int id) : tis_(tis), id_(id) {}
  const TiledIndexSpace& tiled_index_space() const { return tis_; }
  int id() const { return id_; }
  bool operator==(const TiledIndexLabel& o) const { return id_ == o.id_; }

 private:
  TiledIndexSpace tis_;
  int id_ = 0;
};
using IndexLabelVec = std::vector<TiledIndexLabel>;

inline TiledIndexLabel TiledIndexSpace::label(const std::string& name) const {
  return TiledIndexLabel((*this)(name), detail::next_label_id());
}
namespace detail {
template <size_t... I>
auto make_labels(const TiledIndexSpace& t, std::index_sequence<I...>) {
  return std::make_tuple(((void)I, t.label())...);
}
}  // namespace detail
template <size_t N>
auto TiledIndexSpace::labels(const std::string& name) const {
  return detail::make_labels((*this)(name), std::make_index_sequence<N>{});
}

// ExecutionContext (execution_context.hpp:330-365): one GPU, one process.  It carries the
// default execution hardware and the profile table Scheduler::execute(hw, true) writes, in
// TAMM's format (get_profile_header).
class ExecutionContext {
 public:
  ExecutionContext() = default;
  explicit ExecutionContext(ExecutionHW hw) : exhw_(hw) {}
  ExecutionHW exhw() const { return exhw_; }
  void set_exhw(ExecutionHW hw) { exhw_ = hw; }
  bool print() const { return true; }
  std::stringstream& get_profile_data() { return profile_data_; }
  std::string get_profile_header() const {
    std::string h = "ID;Level;OP;total_op_time_min;total_op_time_max;total_op_time_avg;";
    h += "get_time_min;get_time_max;get_time_avg;";
    h += "block_compute_time_min;block_compute_time_max;block_compute_time_avg;";
    h += "copy_time_min;copy_time_max;copy_time_avg;";
    h += "acc_time_min;acc_time_max;acc_time_avg";
    return h;
  }

 private:
  ExecutionHW exhw_ = ExecutionHW::GPU;
  std::stringstream profile_data_;
};

template <typename T>
class LabeledTensor;

template <typename T>
class Tensor {
  static_assert(std::is_same<T, double>::value, "the MI355X engine computes in FP64 (T = double)");

 public:
  Tensor() : impl_(std::make_shared<Impl>()) { make({}, -1, 0); }
  Tensor(std::initializer_list<TiledIndexSpace> spaces) : Tensor(std::vector<TiledIndexSpace>(spaces)) {}
  explicit Tensor(std::vector<TiledIndexSpace> spaces) : impl_(std::make_shared<Impl>()) {
    make(std::move(spaces), -1, 0);
  }
  Tensor(std::vector<TiledIndexSpace> spaces, std::vector<size_t> spin_sizes) : impl_(std::make_shared<Impl>()) {
    int up = spin_sizes.empty() ? -1 : (int)spin_sizes[0];
    int lo = spin_sizes.size() > 1 ? (int)spin_sizes[1] : (int)spaces.size() - up;
    make(std::move(spaces), up, lo);
  }
  Tensor(std::initializer_list<TiledIndexLabel> labels) : impl_(std::make_shared<Impl>()) {
    std::vector<TiledIndexSpace> s;
    for (auto& l : labels) s.push_back(l.tiled_index_space());
    make(std::move(s), -1, 0);
  }

  template <typename... Ls>
  LabeledTensor<T> operator()(const Ls&... ls) const {
    return LabeledTensor<T>(*this, IndexLabelVec{ls...});
  }
  LabeledTensor<T> operator()() const {
    IndexLabelVec v;
    for (auto& s : impl_->spaces) v.push_back(s.label());
    return LabeledTensor<T>(*this, v);
  }

  size_t num_modes() const { return impl_->spaces.size(); }
  const std::vector<TiledIndexSpace>& tiled_index_spaces() const { return impl_->spaces; }
  tmx_tensor handle() const { return impl_->h; }
  double* data() const { return impl_->data; }
  int64_t size() const { return impl_->size; }
  bool is_allocated() const { return impl_->data != nullptr; }

  size_t block_size(const IndexVector& bid) const {
    size_t s = 1;
    for (size_t i = 0; i < bid.size(); ++i) s *= impl_->spaces[i].tile_size(bid[i]);
    return s;
  }
  std::vector<size_t> block_dims(const IndexVector& bid) const {
    std::vector<size_t> d;
    for (size_t i = 0; i < bid.size(); ++i) d.push_back(impl_->spaces[i].tile_size(bid[i]));
    return d;
  }
  bool is_non_zero(const IndexVector& bid) const { return offset_of(bid) >= 0; }
  // Copies one block to / from the host (TensorImpl::get / put, tensor_impl.hpp:321-337).
  void get(const IndexVector& bid, std::vector<T>& buf) const {
    int64_t off = offset_of(bid);
    buf.assign(block_size(bid), T{0});
    if (off >= 0)
      detail::check(tmx_memcpy(buf.data(), impl_->data + off, buf.size() * sizeof(T), TMX_D2H, nullptr));
  }
  void put(const IndexVector& bid, const std::vector<T>& buf) {
    int64_t off = offset_of(bid);
    if (off < 0) throw std::runtime_error("put: block is zero by symmetry");
    detail::check(tmx_memcpy(impl_->data + off, buf.data(), buf.size() * sizeof(T), TMX_H2D, nullptr));
  }

  void allocate() {
    if (!impl_->data) {
      void* p = nullptr;
      detail::check(tmx_malloc(sizeof(T) * (size_t)std::max<int64_t>(impl_->size, 1), &p));
      impl_->data = static_cast<double*>(p);
    }
  }
  void deallocate() {
    if (impl_->data) tmx_free(impl_->data);
    impl_->data = nullptr;
  }
  template <typename... Ts>
  static void allocate(ExecutionContext* /*ec*/, Ts&... ts) {
    (ts.allocate(), ...);
  }
  template <typename... Ts>
  static void deallocate(Ts&... ts) {
    (ts.deallocate(), ...);
  }

 private:
  struct Impl {
    std::vector<TiledIndexSpace> spaces;
    tmx_tensor h = nullptr;
    double* data = nullptr;
    int64_t size = 0;
    ~Impl() {
      if (data) tmx_free(data);
      if (h) tmx_tensor_destroy(h);
    }
  };
  void make(std::vector<TiledIndexSpace> spaces, int up, int lo) {
    impl_->spaces = std::move(spaces);
    std::vector<tmx_space> hs;
    for (auto& s : impl_->spaces) hs.push_back(s.handle());
    detail::check(tmx_tensor_create((int)hs.size(), hs.data(), up, lo, 0, &impl_->h));
    detail::check(tmx_tensor_size(impl_->h, &impl_->size));
  }
  int64_t offset_of(const IndexVector& bid) const {
    std::vector<int64_t> b(bid.begin(), bid.end());
    int64_t off = -1, sz = 0;
    detail::check(tmx_tensor_locate(impl_->h, b.data(), &off, &sz));
    return off;
  }
  std::shared_ptr<Impl> impl_;
};

// ---- operations -------------------------------------------------------------------------
class Op {
 public:
  virtual ~Op() = default;
  virtual void execute(void* stream) = 0;
  virtual std::string name() const = 0;
  // The tensor the op overwrites or accumulates into and the tensors it reads, by engine
  // handle (null: none); the Scheduler levelizes on them (scheduler.hpp:80-91, :140-158).
  virtual const void* writes() const = 0;
  virtual const void* accumulates() const = 0;
  virtual std::vector<const void*> reads() const = 0;
  std::string opstr_;                          // label given to Scheduler::operator()
  ExecutionHW exhw_ = ExecutionHW::DEFAULT;    // per-op hardware (op_base.hpp:28)
};

namespace detail {
// Engine operand of a labeled tensor: label ids plus each label's tile range translated into
// the tensor's mode space (internal::translate_blockid, src/tamm/utils.hpp:263-286).
struct OperandData {
  std::vector<int> labels;
  std::vector<int64_t> toff, ntl;
  tmx_operand op{};
};
template <typename T>
std::unique_ptr<OperandData> operand(const Tensor<T>& t, const IndexLabelVec& ls) {
  auto d = std::make_unique<OperandData>();
  if (ls.size() != t.num_modes()) throw std::runtime_error("label count does not match tensor rank");
  for (size_t i = 0; i < ls.size(); ++i) {
    const TiledIndexSpace& ms = t.tiled_index_spaces()[i];
    const TiledIndexSpace& lsp = ls[i].tiled_index_space();
    if (!ms.same_root(lsp)) throw std::runtime_error("label is not over the tensor's index space");
    int64_t off = lsp.tile_offset() - ms.tile_offset();
    if (off < 0 || off + lsp.num_tiles() > ms.num_tiles())
      throw std::runtime_error("label space is not a subspace of the tensor's mode space");
    d->labels.push_back(ls[i].id());
    d->toff.push_back(off);
    d->ntl.push_back(lsp.num_tiles());
  }
  d->op.tensor = t.handle();
  d->op.labels = d->labels.data();
  d->op.tile_offset = d->toff.data();
  d->op.label_ntiles = d->ntl.data();
  return d;
}
}  // namespace detail

template <typename T>
class LabeledTensor;
// Right-hand sides built by operator* (labeled_tensor.hpp:274-290): alpha * A, alpha * A * B.
template <typename T>
struct ScaledExpr {
  T alpha;
  std::shared_ptr<const LabeledTensor<T>> a;
};
template <typename T>
struct MultExpr {
  T alpha;
  std::shared_ptr<const LabeledTensor<T>> a, b;
};

template <typename T>
class LabeledTensor {
 public:
  LabeledTensor(Tensor<T> t, IndexLabelVec ls) : t_(std::move(t)), ls_(std::move(ls)) {}
  const Tensor<T>& tensor() const { return t_; }
  const IndexLabelVec& labels() const { return ls_; }

  std::shared_ptr<Op> operator=(T v) const;  // SetOp
  std::shared_ptr<Op> operator=(const MultExpr<T>& r) const;
  std::shared_ptr<Op> operator+=(const MultExpr<T>& r) const;
  std::shared_ptr<Op> operator-=(const MultExpr<T>& r) const { return *this += MultExpr<T>{-r.alpha, r.a, r.b}; }
  std::shared_ptr<Op> operator=(const ScaledExpr<T>& r) const;  // AddOp (assign)
  std::shared_ptr<Op> operator+=(const ScaledExpr<T>& r) const;
  std::shared_ptr<Op> operator+=(const LabeledTensor& r) const;
  std::shared_ptr<Op> operator=(const LabeledTensor& r) const;

 private:
  Tensor<T> t_;
  IndexLabelVec ls_;
};

// alpha * A, A * B, alpha * A * B (labeled_tensor.hpp:274-290)
template <typename T>
ScaledExpr<T> operator*(T a, const LabeledTensor<T>& x) {
  return {a, std::make_shared<const LabeledTensor<T>>(x)};
}
template <typename T>
MultExpr<T> operator*(const ScaledExpr<T>& ax, const LabeledTensor<T>& y) {
  return {ax.alpha, ax.a, std::make_shared<const LabeledTensor<T>>(y)};
}
template <typename T>
MultExpr<T> operator*(const LabeledTensor<T>& x, const LabeledTensor<T>& y) {
  return {T{1}, std::make_shared<const LabeledTensor<T>>(x), std::make_shared<const LabeledTensor<T>>(y)};
}

template <typename T>
class SetOp : public Op {
 public:
  SetOp(LabeledTensor<T> lhs, T v) : lhs_(std::move(lhs)), v_(v) {}
  void execute(void* stream) override {
    const Tensor<T>& t = lhs_.tensor();
    // Whole tensor when every label spans its mode; otherwise the addressed blocks only.
    bool whole = true;
    for (size_t i = 0; i < t.num_modes(); ++i)
      if (lhs_.labels()[i].tiled_index_space().num_tiles() != t.tiled_index_spaces()[i].num_tiles()) whole = false;
    if (whole) {
      detail::check(tmx_fill(t.data(), t.size(), v_, stream));
      return;
    }
    auto od = detail::operand(t, lhs_.labels());
    std::vector<int64_t> it(t.num_modes(), 0), bid(t.num_modes());
    while (true) {
      for (size_t i = 0; i < it.size(); ++i) bid[i] = od->toff[i] + it[i];
      int64_t off = -1, sz = 0;
      detail::check(tmx_tensor_locate(t.handle(), bid.data(), &off, &sz));
      if (off >= 0) detail::check(tmx_fill(t.data() + off, sz, v_, stream));
      int d = (int)it.size() - 1;
      for (; d >= 0; --d) {
        if (++it[d] < od->ntl[d]) break;
        it[d] = 0;
      }
      if (d < 0) break;
    }
  }
  std::string name() const override { return "SetOp"; }
  const void* writes() const override { return lhs_.tensor().handle(); }
  const void* accumulates() const override { return nullptr; }
  std::vector<const void*> reads() const override { return {}; }

 private:
  LabeledTensor<T> lhs_;
  T v_;
};

template <typename T>
class MultOp : public Op {
 public:
  // MultOp ctor and validate (multop.hpp:207-239, :986-1021).  `=` is TAMM's zero fill of the
  // addressed C blocks plus `+=` (canonicalize, multop.hpp:255-269); the engine fuses the fill
  // into the contraction (tmx_contraction_set_assign).
  MultOp(LabeledTensor<T> c, T alpha, LabeledTensor<T> a, LabeledTensor<T> b, bool is_assign)
      : c_(std::move(c)), a_(std::move(a)), b_(std::move(b)), alpha_(alpha), assign_(is_assign) {
    if (c_.tensor().handle() == a_.tensor().handle() || c_.tensor().handle() == b_.tensor().handle())
      throw std::runtime_error("MultOp: self assignment is not supported");
  }
  ~MultOp() override {
    if (plan_) tmx_contraction_destroy(plan_);
  }
  void set_mode(int mode) { mode_ = mode; }
  void execute(void* stream) override {
    if (!plan_) {
      auto oc = detail::operand(c_.tensor(), c_.labels());
      auto oa = detail::operand(a_.tensor(), a_.labels());
      auto ob = detail::operand(b_.tensor(), b_.labels());
      detail::check(tmx_contraction_create(&oc->op, alpha_, &oa->op, &ob->op, mode_, &plan_));
      if (assign_) detail::check(tmx_contraction_set_assign(plan_, 1));
    }
    detail::check(tmx_contraction_execute(plan_, a_.tensor().data(), b_.tensor().data(), c_.tensor().data(), stream));
  }
  std::string name() const override { return "MultOp"; }
  const void* writes() const override { return assign_ ? c_.tensor().handle() : nullptr; }
  const void* accumulates() const override { return assign_ ? nullptr : c_.tensor().handle(); }
  std::vector<const void*> reads() const override { return {a_.tensor().handle(), b_.tensor().handle()}; }

 private:
  LabeledTensor<T> c_, a_, b_;
  T alpha_;
  bool assign_;
  int mode_ = TMX_EXEC_AUTO;
  tmx_contraction plan_ = nullptr;
};

template <typename T>
class AddOp : public Op {
 public:
  AddOp(LabeledTensor<T> c, T alpha, LabeledTensor<T> a, bool is_assign)
      : c_(std::move(c)), a_(std::move(a)), alpha_(alpha), assign_(is_assign) {}
  ~AddOp() override {
    if (plan_) tmx_addop_destroy(plan_);
  }
  void execute(void* stream) override {
    if (!plan_) {
      auto oc = detail::operand(c_.tensor(), c_.labels());
      auto oa = detail::operand(a_.tensor(), a_.labels());
      detail::check(tmx_addop_create(&oc->op, alpha_, &oa->op, assign_ ? 1 : 0, &plan_));
    }
    detail::check(tmx_addop_execute(plan_, a_.tensor().data(), c_.tensor().data(), stream));
  }
  std::string name() const override { return "AddOp"; }
  const void* writes() const override { return assign_ ? c_.tensor().handle() : nullptr; }
  const void* accumulates() const override { return assign_ ? nullptr : c_.tensor().handle(); }
  std::vector<const void*> reads() const override { return {a_.tensor().handle()}; }

 private:
  LabeledTensor<T> c_, a_;
  T alpha_;
  bool assign_;
  tmx_addop plan_ = nullptr;
};

template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator=(T v) const {
  return std::make_shared<SetOp<T>>(*this, v);
}
template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator=(const MultExpr<T>& r) const {
  return std::make_shared<MultOp<T>>(*this, r.alpha, *r.a, *r.b, true);
}
template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator+=(const MultExpr<T>& r) const {
  return std::make_shared<MultOp<T>>(*this, r.alpha, *r.a, *r.b, false);
}
template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator=(const ScaledExpr<T>& r) const {
  return std::make_shared<AddOp<T>>(*this, r.alpha, *r.a, true);
}
template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator+=(const ScaledExpr<T>& r) const {
  return std::make_shared<AddOp<T>>(*this, r.alpha, *r.a, false);
}
template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator+=(const LabeledTensor& r) const {
  return std::make_shared<AddOp<T>>(*this, T{1}, r, false);
}
template <typename T>
std::shared_ptr<Op> LabeledTensor<T>::operator=(const LabeledTensor& r) const {
  return std::make_shared<AddOp<T>>(*this, T{1}, r, true);
}

// Adjoint contractions of C(c) += alpha * A(a) * B(b).  Given dC(c) = dL/dC (labeled with C's
// labels), returns dA(a) += alpha * dC(c) * B(b) and dB(b) += alpha * dC(c) * A(a): labeled
// block-sparse contractions of the same kind, run by the same engine kernels, with dA / dB
// sharing A / B's block structure (so the gradient has the operand's sparsity).  A null
// dA or dB treats that operand as a constant.  Queue the ops on a Scheduler like any other.
template <typename T>
std::vector<std::shared_ptr<Op>> adjoint(const LabeledTensor<T>& dC, const MultExpr<T>& rhs,
                                         const Tensor<T>* dA, const Tensor<T>* dB) {
  std::vector<std::shared_ptr<Op>> ops;
  if (dA)
    ops.push_back(std::make_shared<MultOp<T>>(LabeledTensor<T>(*dA, rhs.a->labels()), rhs.alpha, dC, *rhs.b, false));
  if (dB)
    ops.push_back(std::make_shared<MultOp<T>>(LabeledTensor<T>(*dB, rhs.b->labels()), rhs.alpha, dC, *rhs.a, false));
  return ops;
}

// A captured Scheduler level: the ops recorded once into a hipGraph and replayed by launch()
// (the op objects, hence their plans and workspaces, live as long as the graph).
class Graph {
 public:
  Graph(std::vector<std::shared_ptr<Op>> ops, void* exec) : ops_(std::move(ops)), exec_(exec) {}
  Graph(const Graph&) = delete;
  Graph& operator=(const Graph&) = delete;
  ~Graph() {
    if (exec_) tmx_graph_destroy(exec_);
  }
  // Replays the level on the engine stream and waits for it.
  void launch() {
    detail::check(tmx_graph_launch(exec_, nullptr));
    detail::check(tmx_synchronize());
  }

 private:
  std::vector<std::shared_ptr<Op>> ops_;
  void* exec_;
};

// Scheduler (scheduler.hpp:40-51, :160-438).  Ops are levelized as TAMM does
// (levelize_and_order, scheduler.hpp:140-158: an op goes one level above the latest earlier op
// it depends on through a read/write, write/write or read-or-write/accumulate overlap; ops
// accumulating into the same tensor share a level) and run level by level, in submission
// order within a level, on the engine's in-order stream.  An op's own ExecutionHW (given to
// operator()) overrides the one passed to execute(), as scheduler.hpp:249-250 does; CPU is
// refused.  With profile=true each op's GPU time goes into ec.get_profile_data() as one
// line of TAMM's profile table (ExecutionContext::get_profile_header, scheduler.hpp:330-396):
// one process, so min = max = avg; tensors are HBM-resident, so the get, copy and acc
// columns are 0 and the op's whole time is its block-compute time.  Times are in seconds.
class Scheduler {
 public:
  explicit Scheduler(ExecutionContext& ec) : ec_(&ec) {}
  Scheduler& operator()(std::shared_ptr<Op> op, std::string opstr = "", ExecutionHW hw = ExecutionHW::DEFAULT) {
    op->opstr_ = std::move(opstr);
    op->exhw_ = hw;
    ops_.push_back(std::move(op));
    return *this;
  }
  template <typename... Ts>
  Scheduler& allocate(Ts&... ts) {
    (ts.allocate(), ...);
    return *this;
  }
  template <typename... Ts>
  Scheduler& deallocate(Ts&... ts) {
    dealloc_.push_back([&ts...]() { (ts.deallocate(), ...); });
    return *this;
  }
  // (level, index into the queued ops), sorted by level, stable.
  std::vector<std::pair<size_t, size_t>> levelize_and_order() const {
    std::vector<std::pair<size_t, size_t>> order;
    for (size_t i = 0; i < ops_.size(); ++i) order.push_back({0, i});
    for (size_t i = 0; i + 1 < ops_.size(); ++i)
      for (size_t j = i + 1; j < ops_.size(); ++j)
        if (depends(*ops_[i], *ops_[j])) order[j].first = std::max(order[i].first + 1, order[j].first);
    std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    return order;
  }
  void execute(ExecutionHW hw = ExecutionHW::DEFAULT, bool profile = false) {
    const auto order = levelize_and_order();
    ExecutionHW execute_on = hw == ExecutionHW::DEFAULT ? ec_->exhw() : hw;
    for (auto& lo : order) {
      const Op& op = *ops_[lo.second];
      if (op.exhw_ != ExecutionHW::DEFAULT) execute_on = op.exhw_;
      if (execute_on == ExecutionHW::CPU)
        throw std::runtime_error("this engine executes on the GPU; use TAMM's CPU Scheduler for ExecutionHW::CPU");
    }
    void* ev0 = nullptr;
    void* ev1 = nullptr;
    if (profile) {
      detail::check(tmx_event_create(&ev0));
      detail::check(tmx_event_create(&ev1));
    }
    int id = 0;
    for (auto& lo : order) {
      Op& op = *ops_[lo.second];
      if (profile) detail::check(tmx_event_record(ev0, nullptr));
      op.execute(nullptr);
      if (profile) {
        detail::check(tmx_event_record(ev1, nullptr));
        float ms = 0;
        detail::check(tmx_event_elapsed_ms(ev0, ev1, &ms));
        profile_.push_back({op.name() + (op.opstr_.empty() ? "" : ":" + op.opstr_), ms});
        const double t = ms * 1e-3;
        ec_->get_profile_data() << id << ";" << lo.first << ";" << op.opstr_ << ";" << t << ";" << t << ";" << t
                                << ";0;0;0;" << t << ";" << t << ";" << t << ";0;0;0;0;0;0" << std::endl;
        total_s_ += t;
      }
      ++id;
    }
    detail::check(tmx_synchronize());
    if (profile) {
      tmx_event_destroy(ev0);
      tmx_event_destroy(ev1);
      ec_->get_profile_data() << ";SUM;;;;" << total_s_ << ";;;0;;;" << total_s_ << ";;;0;;;0" << std::endl;
      total_s_ = 0;
    }
    for (auto& f : dealloc_) f();
    ops_.clear();
    dealloc_.clear();
  }
  // Executes the queued ops once (building every plan and block table), then records the
  // same ops into a hipGraph for replay: a Scheduler level as one launch.  The eager pass
  // has the ops' effects like execute(); each launch() of the result repeats them.
  std::shared_ptr<Graph> capture() {
    std::vector<std::shared_ptr<Op>> ops;
    for (auto& lo : levelize_and_order()) {
      if (ops_[lo.second]->exhw_ == ExecutionHW::CPU)
        throw std::runtime_error("this engine executes on the GPU; use TAMM's CPU Scheduler for ExecutionHW::CPU");
      ops_[lo.second]->execute(nullptr);
      ops.push_back(ops_[lo.second]);
    }
    detail::check(tmx_synchronize());
    detail::check(tmx_graph_begin(nullptr));
    int rc = TMX_OK;
    try {
      for (auto& op : ops) op->execute(nullptr);
    } catch (...) {
      void* g = nullptr;
      tmx_graph_end(nullptr, &g);
      if (g) tmx_graph_destroy(g);
      throw;
    }
    void* exec = nullptr;
    rc = tmx_graph_end(nullptr, &exec);
    detail::check(rc);
    for (auto& f : dealloc_) f();
    ops_.clear();
    dealloc_.clear();
    return std::make_shared<Graph>(std::move(ops), exec);
  }
  const std::vector<std::pair<std::string, float>>& profile_data() const { return profile_; }
  ExecutionContext& ec() { return *ec_; }

 private:
  static bool overlap(const std::vector<const void*>& a, const void* t) {
    return t && std::find(a.begin(), a.end(), t) != a.end();
  }
  // has_dependence (scheduler.hpp:85-91): R1/W2, W1/R2, W1/W2, R1/A2, W1/A2, R2/A1, W2/A1.
  static bool depends(const Op& x, const Op& y) {
    const auto rx = x.reads(), ry = y.reads();
    const void *wx = x.writes(), *wy = y.writes(), *ax = x.accumulates(), *ay = y.accumulates();
    return overlap(rx, wy) || overlap(ry, wx) || (wx && wx == wy) || overlap(rx, ay) ||
           (wx && wx == ay) || overlap(ry, ax) || (wy && wy == ax);
  }
  ExecutionContext* ec_;
  std::vector<std::shared_ptr<Op>> ops_;
  std::vector<std::function<void()>> dealloc_;
  std::vector<std::pair<std::string, float>> profile_;
  double total_s_ = 0;
};

// write_to_disk / read_from_disk (tamm_io.hpp:223-460, :735-980; the tammio format): one HDF5
// dataset "tensor" of the tensor's dense extent holding the nonzero blocks back to back in
// row-major block order.  Synchronous.
template <typename T>
void write_to_disk(const Tensor<T>& t, const std::string& filename) {
  detail::check(tmx_tensor_write_hdf5(t.handle(), t.data(), filename.c_str(), nullptr));
}
template <typename T>
void read_from_disk(Tensor<T>& t, const std::string& filename) {
  detail::check(tmx_tensor_read_hdf5(t.handle(), t.data(), filename.c_str(), nullptr));
}

// tamm::norm (tamm_utils.hpp:1197-1254): sqrt of the sum of squares of the stored blocks.
template <typename T>
T norm(const Tensor<T>& t) {
  double r = 0;
  detail::check(tmx_norm(t.data(), t.size(), &r, nullptr));
  return r;
}

namespace kernels {
namespace detail {
// The stream of TAMM's gpuStream_t (std::pair<hipStream_t, rocblas_handle>,
// gpu_streams.hpp:31) or of a plain stream / null (the engine stream).
template <typename S, typename H>
void* stream_of(std::pair<S, H>& h) {
  return reinterpret_cast<void*>(h.first);
}
inline void* stream_of(void* s) { return s; }
inline void* stream_of(std::nullptr_t) { return nullptr; }
// An extent of TAMM's SizeVec: Size (StrongNum<OffsetSpace, uint64_t>, types.hpp:73-89)
// through value(), or any integer.
template <typename X>
auto extent_of(const X& x, int) -> decltype((int64_t)x.value()) {
  return (int64_t)x.value();
}
template <typename X>
int64_t extent_of(const X& x, long) {
  return (int64_t)x;
}
template <typename V>
std::vector<int64_t> extents(const V& v) {
  std::vector<int64_t> out;
  for (const auto& x : v) out.push_back(extent_of(x, 0));
  return out;
}
template <typename V>
std::vector<int> label_ints(const V& v) {
  return std::vector<int>(v.begin(), v.end());
}
}  // namespace detail

// kernels::block_multiply with TAMM's GPU-build signature (multiply.hpp:212-220), called as
// execute_bufacc and Test_CCSD do (multop.hpp:793-798, Test_CCSD.cpp:436-441).  `thandle` is
// TAMM's gpuStream_t (its stream is used) or a plain stream; the dims are TAMM's SizeVec or
// any vector of integers.  Host abuf / bbuf are copied into th_a / th_b (device buffers of at
// least asize / bsize elements the caller owns, multop.hpp:785-790), then one fused
// permute -> FP64 MFMA -> permute kernel writes C-order results into cinter_buf_dev:
//   beta == 0: cinter_buf_dev = alpha * perm(A*B)
//   else:      cinter_buf_dev = alpha * perm(A*B) + beta * cinter_buf_dev.
// TAMM forms alpha*A*B + beta*X in its GEMM-order buffer cinter_tmp_buf_dev and permutes that
// into cinter_buf_dev (multiply.hpp:42-74, :196-210), so when cinter_buf_dev holds the
// permuted copy of that buffer (both zeroed by the caller, as execute_bufacc does) the results
// are the same; cinter_tmp_buf_dev and cbuf are not touched.  Labels reduced on one side only
// are summed over, at any beta (TAMM's per-slice GEMMs overwrite one another when beta = 0).
template <typename T, typename T1, typename T2, typename T3, typename Handle, typename AV, typename BV,
          typename CV, typename AL, typename BL, typename CL>
void block_multiply(T2*& th_a, T3*& th_b, Handle& thandle, T alpha, const T2* abuf, const AV& adims,
                    const AL& alabels, const T3* bbuf, const BV& bdims, const BL& blabels, T beta,
                    T1* /*cbuf*/, const CV& cdims, const CL& clabels, ExecutionHW hw, bool /*is_assign*/,
                    T1*& cinter_buf_dev, T1*& /*cinter_tmp_buf_dev*/) {
  static_assert(std::is_same<T1, double>::value && std::is_same<T2, double>::value && std::is_same<T3, double>::value,
                "real FP64 blocks only");
  if (hw == ExecutionHW::CPU) throw std::runtime_error("block_multiply: CPU execution is TAMM's own path");
  void* s = detail::stream_of(thandle);
  const std::vector<int64_t> ad = detail::extents(adims), bd = detail::extents(bdims), cd = detail::extents(cdims);
  const std::vector<int> al = detail::label_ints(alabels), bl = detail::label_ints(blabels),
                         cl = detail::label_ints(clabels);
  int64_t asz = 1, bsz = 1, csz = 1;
  for (auto d : ad) asz *= d;
  for (auto d : bd) bsz *= d;
  for (auto d : cd) csz *= d;
  ::tamm::detail::check(tmx_memcpy(th_a, abuf, (size_t)asz * sizeof(T2), TMX_H2D, s));
  ::tamm::detail::check(tmx_memcpy(th_b, bbuf, (size_t)bsz * sizeof(T3), TMX_H2D, s));
  const double b = static_cast<double>(beta);
  if (b != 0.0 && b != 1.0) ::tamm::detail::check(tmx_scale(cinter_buf_dev, csz, b, s));
  ::tamm::detail::check(tmx_block_multiply(static_cast<double>(alpha), th_a, (int)ad.size(), ad.data(), al.data(),
                                           th_b, (int)bd.size(), bd.data(), bl.data(), cinter_buf_dev,
                                           (int)cd.size(), cd.data(), cl.data(), b == 0.0 ? 1 : 0, s));
}
}  // namespace kernels

}  // namespace tamm
