// The cudf C++ API for the hash groupby path (cudf::groupby::groupby,
// aggregation factories, cudf::hashing row hashers) implemented over the C
// ABI of mgb_groupby.h. Argument checks and their exceptions follow
// cpp/src/groupby/groupby.cu:173-219; results are columns whose buffers come
// from the caller's memory resource.
#include <cudf/aggregation.hpp>
#include <cudf/groupby.hpp>
#include <cudf/hashing.hpp>
#include <cudf/utilities/error.hpp>

#include <rmm/device_buffer.hpp>

#include <hip/hip_runtime.h>

#include "mgb_groupby.h"

#include <string>
#include <vector>

namespace cudf {

std::size_t size_of(data_type t)
{
  switch (t.id()) {
    case type_id::INT8: case type_id::UINT8: case type_id::BOOL8: return 1;
    case type_id::INT16: case type_id::UINT16: return 2;
    case type_id::INT32: case type_id::UINT32: case type_id::FLOAT32: case type_id::TIMESTAMP_DAYS:
    case type_id::DURATION_DAYS: case type_id::DECIMAL32: return 4;
    case type_id::INT64: case type_id::UINT64: case type_id::FLOAT64: case type_id::TIMESTAMP_SECONDS:
    case type_id::TIMESTAMP_MILLISECONDS: case type_id::TIMESTAMP_MICROSECONDS: case type_id::TIMESTAMP_NANOSECONDS:
    case type_id::DURATION_SECONDS: case type_id::DURATION_MILLISECONDS: case type_id::DURATION_MICROSECONDS:
    case type_id::DURATION_NANOSECONDS: case type_id::DECIMAL64: return 8;
    case type_id::DECIMAL128: return 16;
    default: CUDF_FAIL("size_of: not a fixed-width type");
  }
}

rmm::cuda_stream_view const get_default_stream() { return rmm::cuda_stream_view{}; }

namespace detail {

// One concrete aggregation type serving every factory; `param` carries ddof
// (VARIANCE / STD) or n (NTH_ELEMENT), `nulls` the null_policy of NUNIQUE /
// NTH_ELEMENT / COLLECT_LIST / COLLECT_SET, `quantiles` / `interp` those of
// QUANTILE, `nulls_equal` / `nans_equal` those of COLLECT_SET.
class simple_aggregation final : public groupby_aggregation,
                                 public groupby_scan_aggregation,
                                 public reduce_aggregation,
                                 public rolling_aggregation {
 public:
  explicit simple_aggregation(aggregation::Kind k, size_type param = 0, null_policy nulls = null_policy::EXCLUDE,
                              std::vector<double> quantiles = {}, interpolation interp = interpolation::LINEAR,
                              null_equality nulls_equal = null_equality::EQUAL,
                              nan_equality nans_equal = nan_equality::ALL_EQUAL)
    : aggregation(k), _param{param}, _nulls{nulls}, _quantiles{std::move(quantiles)}, _interp{interp},
      _nulls_equal{nulls_equal}, _nans_equal{nans_equal}
  {
  }
  [[nodiscard]] std::unique_ptr<aggregation> clone() const override
  {
    return std::make_unique<simple_aggregation>(kind, _param, _nulls, _quantiles, _interp, _nulls_equal,
                                                _nans_equal);
  }
  [[nodiscard]] bool is_equal(aggregation const& other) const override
  {
    auto const* o = dynamic_cast<simple_aggregation const*>(&other);
    return o && o->kind == kind && o->_param == _param && o->_nulls == _nulls && o->_quantiles == _quantiles &&
           o->_interp == _interp && o->_nulls_equal == _nulls_equal && o->_nans_equal == _nans_equal;
  }
  [[nodiscard]] size_type param() const { return _param; }
  [[nodiscard]] null_policy nulls() const { return _nulls; }
  [[nodiscard]] std::vector<double> const& quantiles() const { return _quantiles; }
  [[nodiscard]] interpolation interp() const { return _interp; }
  [[nodiscard]] null_equality nulls_equal() const { return _nulls_equal; }
  [[nodiscard]] nan_equality nans_equal() const { return _nans_equal; }

 private:
  size_type _param;
  null_policy _nulls;
  std::vector<double> _quantiles;
  interpolation _interp;
  null_equality _nulls_equal;
  nan_equality _nans_equal;
};

// The C ABI aggregations of one cudf aggregation: one per quantile of a
// QUANTILE, one otherwise.
std::vector<mgb_aggregation> to_mgb_aggs(aggregation const& a)
{
  auto const* s = dynamic_cast<simple_aggregation const*>(&a);
  mgb_aggregation m{};
  m.kind = static_cast<int32_t>(a.kind);
  if (!s) return {m};
  m.ddof = s->param();
  m.null_policy = s->nulls() == null_policy::INCLUDE ? MGB_NULL_INCLUDE : MGB_NULL_EXCLUDE;
  m.interpolation = static_cast<int32_t>(s->interp());
  if (a.kind == aggregation::COLLECT_SET) {
    m.ddof = s->nulls_equal() == null_equality::EQUAL ? MGB_NULL_EQUAL : MGB_NULL_UNEQUAL;
    m.interpolation = s->nans_equal() == nan_equality::ALL_EQUAL ? MGB_NAN_ALL_EQUAL : MGB_NAN_UNEQUAL;
  }
  if (a.kind != aggregation::QUANTILE) return {m};
  std::vector<mgb_aggregation> out;
  for (double q : s->quantiles()) {
    m.quantile = q;
    out.push_back(m);
  }
  return out;
}

// Validity of a flattened [groups x k] quantile result: bit g * k + j is bit g
// of the groups' validity.
__global__ void repeat_bits_kernel(const uint32_t* in, int64_t groups, int32_t k, uint32_t* out)
{
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t words = (groups * k + 31) / 32;
  if (w >= words) return;
  uint32_t v = 0;
  for (int b = 0; b < 32; ++b) {
    const int64_t i = w * 32 + b;
    if (i >= groups * k) break;
    const int64_t g = i / k;
    v |= ((in[g >> 5] >> (g & 31)) & 1u) << b;
  }
  out[w] = v;
}

[[noreturn]] void throw_status(mgb_status st)
{
  std::string msg = mgb_last_error_message();
  switch (st) {
    case MGB_ERROR_HIP: throw cudf::cuda_error("CUDA error: " + msg);
    case MGB_ERROR_ALLOC: throw rmm::bad_alloc(msg);
    default: throw cudf::logic_error(msg);
  }
}

// Bridges mgb_allocator to an rmm device_memory_resource. Exceptions from the
// resource are turned into a NULL return (the C side reports MGB_ERROR_ALLOC).
struct mr_allocator {
  static void* allocate(void* ctx, size_t bytes, void* stream)
  {
    try {
      return static_cast<rmm::mr::device_memory_resource*>(ctx)->allocate(
        bytes, rmm::cuda_stream_view{static_cast<hipStream_t>(stream)});
    } catch (...) {
      return nullptr;
    }
  }
  static void deallocate(void* ctx, void* p, size_t bytes, void* stream)
  {
    static_cast<rmm::mr::device_memory_resource*>(ctx)->deallocate(
      p, bytes, rmm::cuda_stream_view{static_cast<hipStream_t>(stream)});
  }
  static mgb_allocator make(rmm::mr::device_memory_resource* mr)
  {
    return mgb_allocator{&mr_allocator::allocate, &mr_allocator::deallocate, mr, nullptr, nullptr};
  }
};

mgb_column to_mgb(column_view const& c)
{
  mgb_column m{};
  m.type_id = static_cast<int32_t>(c.type().id());
  m.size = c.size();
  m.data = c.head();
  m.null_mask = c.null_mask();
  m.offset = c.offset();
  m.null_count = c.nullable() ? c.null_count() : 0;
  return m;
}

std::unique_ptr<column> adopt(mgb_out_column const& o, rmm::cuda_stream_view stream,
                              rmm::mr::device_memory_resource* mr)
{
  rmm::device_buffer data{o.data, o.data ? o.data_bytes : 0, stream, mr};
  rmm::device_buffer mask{o.null_mask, o.null_mask ? o.mask_bytes : 0, stream, mr};
  if (o.type_id == MGB_TYPE_LIST) {
    // cudf's lists column: no data of its own, children {INT32 offsets, values}
    // (cpp/include/cudf/lists/lists_column_view.hpp:63-64)
    std::vector<std::unique_ptr<column>> children;
    children.push_back(std::make_unique<column>(data_type{type_id::INT32},
                                                static_cast<size_type>(o.size ? o.size + 1 : 0), std::move(data),
                                                rmm::device_buffer{}, 0));
    rmm::device_buffer cdata{o.child_data, o.child_data ? o.child_data_bytes : 0, stream, mr};
    rmm::device_buffer cmask{o.child_null_mask, o.child_null_mask ? o.child_mask_bytes : 0, stream, mr};
    children.push_back(std::make_unique<column>(data_type{static_cast<type_id>(o.child_type_id)},
                                                static_cast<size_type>(o.child_size), std::move(cdata),
                                                std::move(cmask), static_cast<size_type>(o.child_null_count)));
    return std::make_unique<column>(data_type{type_id::LIST}, static_cast<size_type>(o.size), rmm::device_buffer{},
                                    std::move(mask), static_cast<size_type>(o.null_count), std::move(children));
  }
  return std::make_unique<column>(data_type{static_cast<type_id>(o.type_id)}, static_cast<size_type>(o.size),
                                  std::move(data), std::move(mask), static_cast<size_type>(o.null_count));
}

}  // namespace detail

#define MGB_FACTORY(name, KIND)                                                   \
  template <typename Base>                                                        \
  std::unique_ptr<Base> name()                                                    \
  {                                                                               \
    return std::make_unique<detail::simple_aggregation>(aggregation::KIND);       \
  }                                                                               \
  template std::unique_ptr<aggregation> name<aggregation>();                      \
  template std::unique_ptr<groupby_aggregation> name<groupby_aggregation>();      \
  template std::unique_ptr<reduce_aggregation> name<reduce_aggregation>();        \
  template std::unique_ptr<rolling_aggregation> name<rolling_aggregation>();

MGB_FACTORY(make_sum_aggregation, SUM)
MGB_FACTORY(make_product_aggregation, PRODUCT)
MGB_FACTORY(make_min_aggregation, MIN)
MGB_FACTORY(make_max_aggregation, MAX)
MGB_FACTORY(make_sum_of_squares_aggregation, SUM_OF_SQUARES)
MGB_FACTORY(make_mean_aggregation, MEAN)
MGB_FACTORY(make_m2_aggregation, M2)
MGB_FACTORY(make_median_aggregation, MEDIAN)
MGB_FACTORY(make_argmax_aggregation, ARGMAX)
MGB_FACTORY(make_argmin_aggregation, ARGMIN)
#undef MGB_FACTORY

// COUNT(EXCLUDE) -> COUNT_VALID, COUNT(INCLUDE) -> COUNT_ALL
// (cpp/src/aggregation/aggregation.cpp:471-476)
template <typename Base>
std::unique_ptr<Base> make_count_aggregation(null_policy null_handling)
{
  return std::make_unique<detail::simple_aggregation>(
    null_handling == null_policy::INCLUDE ? aggregation::COUNT_ALL : aggregation::COUNT_VALID);
}
template std::unique_ptr<aggregation> make_count_aggregation<aggregation>(null_policy);
template std::unique_ptr<groupby_aggregation> make_count_aggregation<groupby_aggregation>(null_policy);
template std::unique_ptr<reduce_aggregation> make_count_aggregation<reduce_aggregation>(null_policy);
template std::unique_ptr<rolling_aggregation> make_count_aggregation<rolling_aggregation>(null_policy);

template <typename Base>
std::unique_ptr<Base> make_variance_aggregation(size_type ddof)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::VARIANCE, ddof);
}
template std::unique_ptr<aggregation> make_variance_aggregation<aggregation>(size_type);
template std::unique_ptr<groupby_aggregation> make_variance_aggregation<groupby_aggregation>(size_type);

template <typename Base>
std::unique_ptr<Base> make_std_aggregation(size_type ddof)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::STD, ddof);
}
template std::unique_ptr<aggregation> make_std_aggregation<aggregation>(size_type);
template std::unique_ptr<groupby_aggregation> make_std_aggregation<groupby_aggregation>(size_type);

template <typename Base>
std::unique_ptr<Base> make_nunique_aggregation(null_policy null_handling)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::NUNIQUE, 0, null_handling);
}
template std::unique_ptr<aggregation> make_nunique_aggregation<aggregation>(null_policy);
template std::unique_ptr<groupby_aggregation> make_nunique_aggregation<groupby_aggregation>(null_policy);

template <typename Base>
std::unique_ptr<Base> make_nth_element_aggregation(size_type n, null_policy null_handling)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::NTH_ELEMENT, n, null_handling);
}
template std::unique_ptr<aggregation> make_nth_element_aggregation<aggregation>(size_type, null_policy);
template std::unique_ptr<groupby_aggregation> make_nth_element_aggregation<groupby_aggregation>(size_type,
                                                                                                null_policy);

template <typename Base>
std::unique_ptr<Base> make_collect_list_aggregation(null_policy null_handling)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::COLLECT_LIST, 0, null_handling);
}
template std::unique_ptr<aggregation> make_collect_list_aggregation<aggregation>(null_policy);
template std::unique_ptr<groupby_aggregation> make_collect_list_aggregation<groupby_aggregation>(null_policy);

template <typename Base>
std::unique_ptr<Base> make_collect_set_aggregation(null_policy null_handling, null_equality nulls_equal,
                                                   nan_equality nans_equal)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::COLLECT_SET, 0, null_handling,
                                                      std::vector<double>{}, interpolation::LINEAR, nulls_equal,
                                                      nans_equal);
}
template std::unique_ptr<aggregation> make_collect_set_aggregation<aggregation>(null_policy, null_equality,
                                                                               nan_equality);
template std::unique_ptr<groupby_aggregation> make_collect_set_aggregation<groupby_aggregation>(null_policy,
                                                                                               null_equality,
                                                                                               nan_equality);

template <typename Base>
std::unique_ptr<Base> make_quantile_aggregation(std::vector<double> const& quantiles, interpolation interp)
{
  return std::make_unique<detail::simple_aggregation>(aggregation::QUANTILE, 0, null_policy::EXCLUDE, quantiles,
                                                      interp);
}
template std::unique_ptr<aggregation> make_quantile_aggregation<aggregation>(std::vector<double> const&,
                                                                             interpolation);
template std::unique_ptr<groupby_aggregation> make_quantile_aggregation<groupby_aggregation>(
  std::vector<double> const&, interpolation);

namespace groupby {

groupby::groupby(table_view const& keys, null_policy null_handling, sorted keys_are_sorted,
                 std::vector<order> const& column_order, std::vector<null_order> const& null_precedence)
  : _keys{keys},
    _include_null_keys{null_handling},
    _keys_are_sorted{keys_are_sorted},
    _column_order{column_order},
    _null_precedence{null_precedence}
{
}

groupby::~groupby() = default;

std::pair<std::unique_ptr<table>, std::vector<aggregation_result>> groupby::aggregate(
  host_span<aggregation_request const> requests, rmm::mr::device_memory_resource* mr)
{
  return aggregate(requests, cudf::get_default_stream(), mr);
}

std::pair<std::unique_ptr<table>, std::vector<aggregation_result>> groupby::aggregate(
  host_span<aggregation_request const> requests, rmm::cuda_stream_view stream, rmm::mr::device_memory_resource* mr)
{
  // cpp/src/groupby/groupby.cu:208-214
  for (auto const& r : requests)
    CUDF_EXPECTS(r.values.size() == _keys.num_rows(), "Size mismatch between request values and groupby keys.");
  for (auto const& r : requests)
    for (auto const& a : r.aggregations)
      CUDF_EXPECTS(mgb_target_type(static_cast<int32_t>(r.values.type().id()), a->kind) >= 0,
                   "Invalid type/aggregation combination.");

  std::vector<mgb_column> keys;
  for (auto const& c : _keys) keys.push_back(detail::to_mgb(c));
  std::vector<std::vector<mgb_aggregation>> aggs(requests.size());
  std::vector<mgb_request> reqs(requests.size());
  std::size_t n_out = 0;
  // C ABI aggregations of each cudf aggregation (a QUANTILE of k quantiles is k of them)
  std::vector<std::vector<std::size_t>> parts(requests.size());
  for (std::size_t i = 0; i < requests.size(); ++i) {
    for (auto const& a : requests[i].aggregations) {
      auto m = detail::to_mgb_aggs(*a);
      parts[i].push_back(m.size());
      aggs[i].insert(aggs[i].end(), m.begin(), m.end());
    }
    reqs[i].values = detail::to_mgb(requests[i].values);
    reqs[i].aggs = aggs[i].data();
    reqs[i].num_aggs = static_cast<int32_t>(aggs[i].size());
    n_out += aggs[i].size();
  }
  std::vector<mgb_out_column> out_keys(keys.size());
  std::vector<mgb_out_column> out_results(n_out);
  int64_t groups = 0;
  // results from `mr`, temporaries from the current device resource (as
  // cudf's hash groupby takes them, cpp/src/groupby/hash/groupby.cu:404,478,484)
  mgb_allocator temps = detail::mr_allocator::make(rmm::mr::get_current_device_resource());
  mgb_allocator alloc = detail::mr_allocator::make(mr);
  alloc.temporaries = &temps;
  mgb_status st = mgb_groupby_aggregate(keys.data(), static_cast<int32_t>(keys.size()),
                                        _include_null_keys == null_policy::INCLUDE ? MGB_NULL_INCLUDE
                                                                                   : MGB_NULL_EXCLUDE,
                                        reqs.data(), static_cast<int32_t>(reqs.size()), out_keys.data(),
                                        out_results.data(), &groups, stream.value(), &alloc);
  if (st != MGB_SUCCESS) detail::throw_status(st);

  std::vector<std::unique_ptr<column>> key_cols;
  for (auto const& o : out_keys) key_cols.push_back(detail::adopt(o, stream, mr));
  std::vector<aggregation_result> results(requests.size());
  std::size_t j = 0;
  for (std::size_t i = 0; i < requests.size(); ++i) {
    for (std::size_t k : parts[i]) {
      if (k == 1) {
        results[i].results.push_back(detail::adopt(out_results[j++], stream, mr));
        continue;
      }
      // k quantiles: cudf's flattened [groups x k] column (group_quantiles.cu:95-100)
      std::vector<std::unique_ptr<column>> q;
      for (std::size_t t = 0; t < k; ++t) q.push_back(detail::adopt(out_results[j++], stream, mr));
      const int64_t n = groups * static_cast<int64_t>(k);
      rmm::device_buffer data{static_cast<std::size_t>(n) * 8, stream, mr};
      for (std::size_t t = 0; t < k && groups > 0; ++t)
        if (hipMemcpy2DAsync(static_cast<uint8_t*>(data.data()) + t * 8, k * 8, q[t]->view().head(), 8, 8,
                             static_cast<std::size_t>(groups), hipMemcpyDeviceToDevice, stream.value()) != hipSuccess)
          throw cudf::cuda_error("CUDA error: interleaving quantiles");
      rmm::device_buffer mask;
      size_type nulls = 0;
      if (q[0]->null_count() > 0) {
        const int64_t words = (n + 31) / 32;
        mask = rmm::device_buffer{static_cast<std::size_t>(((words * 4 + 63) / 64) * 64), stream, mr};
        detail::repeat_bits_kernel<<<(unsigned)((words + 255) / 256), 256, 0, stream.value()>>>(
          q[0]->view().null_mask(), groups, static_cast<int32_t>(k), static_cast<uint32_t*>(mask.data()));
        if (hipGetLastError() != hipSuccess) throw cudf::cuda_error("CUDA error: quantile validity");
        nulls = q[0]->null_count() * static_cast<size_type>(k);
      }
      results[i].results.push_back(std::make_unique<column>(data_type{type_id::FLOAT64}, static_cast<size_type>(n),
                                                            std::move(data), std::move(mask), nulls));
    }
  }
  return {std::make_unique<table>(std::move(key_cols)), std::move(results)};
}

}  // namespace groupby

namespace hashing {

namespace {
template <typename Fn>
std::unique_ptr<column> hash_table(table_view const& input, uint32_t seed, rmm::cuda_stream_view stream,
                                   rmm::mr::device_memory_resource* mr, type_id out_type, Fn fn)
{
  std::vector<mgb_column> cols;
  for (auto const& c : input) cols.push_back(detail::to_mgb(c));
  size_type rows = input.num_rows();
  rmm::device_buffer data{static_cast<std::size_t>(rows) * 4, stream, mr};
  if (rows > 0) {
    mgb_status st = fn(cols.data(), static_cast<int32_t>(cols.size()), seed, data.data(), stream.value());
    if (st != MGB_SUCCESS) detail::throw_status(st);
  }
  return std::make_unique<column>(data_type{out_type}, rows, std::move(data), rmm::device_buffer{}, 0);
}
}  // namespace

std::unique_ptr<column> murmurhash3_x86_32(table_view const& input, uint32_t seed, rmm::cuda_stream_view stream,
                                           rmm::mr::device_memory_resource* mr)
{
  return hash_table(input, seed, stream, mr, type_id::UINT32,
                    [](mgb_column const* c, int32_t n, uint32_t s, void* out, hipStream_t st) {
                      return mgb_murmurhash3_x86_32(c, n, s, static_cast<uint32_t*>(out), st);
                    });
}

std::unique_ptr<column> spark_murmurhash3_x86_32(table_view const& input, uint32_t seed,
                                                 rmm::cuda_stream_view stream, rmm::mr::device_memory_resource* mr)
{
  return hash_table(input, seed, stream, mr, type_id::INT32,
                    [](mgb_column const* c, int32_t n, uint32_t s, void* out, hipStream_t st) {
                      return mgb_spark_murmurhash3_x86_32(c, n, s, static_cast<int32_t*>(out), st);
                    });
}

}  // namespace hashing
}  // namespace cudf
