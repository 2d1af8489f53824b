// C++ tests of the cudf::groupby drop-in, written against the same API and
// literal fixtures as the reference's cpp/tests/groupby/*_tests.cpp, with the
// reference's comparison protocol (sort both sides by key, then compare;
// cpp/tests/groupby/groupby_test_util.cpp:33-95). Needs a GPU.
#include <cudf/aggregation.hpp>
#include <cudf/groupby.hpp>
#include <cudf/lists/lists_column_view.hpp>
#include <cudf/hashing.hpp>
#include <cudf/utilities/error.hpp>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

static int g_failures = 0;
static int g_checks = 0;
#define CHECK(cond)                                                                 \
  do {                                                                              \
    ++g_checks;                                                                     \
    if (!(cond)) {                                                                  \
      ++g_failures;                                                                 \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);               \
    }                                                                               \
  } while (0)

// ---------------------------------------------------------------- helpers

// A device column built from host values and optional validity (the role of
// cudf::test::fixed_width_column_wrapper).
template <typename T>
struct wrapper {
  std::unique_ptr<cudf::column> col;
  wrapper(cudf::type_id id, std::vector<T> vals, std::vector<int> valid = {})
  {
    auto* mr = rmm::mr::get_current_device_resource();
    const size_t n = vals.size();
    rmm::device_buffer data{n * sizeof(T), cudf::get_default_stream(), mr};
    if (n) (void)hipMemcpy(data.data(), vals.data(), n * sizeof(T), hipMemcpyHostToDevice);
    rmm::device_buffer mask{};
    int nulls = 0;
    if (!valid.empty()) {
      std::vector<uint32_t> words(((n + 31) / 32 + 15) / 16 * 16, 0u);
      for (size_t i = 0; i < n; ++i) {
        if (valid[i]) words[i / 32] |= 1u << (i % 32);
        else ++nulls;
      }
      mask = rmm::device_buffer{words.size() * 4, cudf::get_default_stream(), mr};
      (void)hipMemcpy(mask.data(), words.data(), words.size() * 4, hipMemcpyHostToDevice);
    }
    col = std::make_unique<cudf::column>(cudf::data_type{id}, static_cast<cudf::size_type>(n), std::move(data),
                                         std::move(mask), nulls);
  }
  cudf::column_view view() const { return col->view(); }
};

template <typename T>
std::vector<T> to_host(cudf::column_view const& c)
{
  std::vector<T> out(c.size());
  (void)hipDeviceSynchronize();
  if (c.size()) (void)hipMemcpy(out.data(), c.data<T>(), c.size() * sizeof(T), hipMemcpyDeviceToHost);
  return out;
}

std::vector<int> valid_host(cudf::column_view const& c)
{
  std::vector<int> v(c.size(), 1);
  if (!c.nullable() || c.size() == 0) return v;
  std::vector<uint32_t> words((c.size() + 31) / 32);
  (void)hipMemcpy(words.data(), c.null_mask(), words.size() * 4, hipMemcpyDeviceToHost);
  for (int i = 0; i < c.size(); ++i) v[i] = (words[(i + c.offset()) / 32] >> ((i + c.offset()) % 32)) & 1;
  return v;
}

// Runs one aggregation and returns (key, value, valid) sorted by key.
template <typename K, typename R>
std::vector<std::tuple<K, R, int>> run_single(cudf::column_view keys, cudf::column_view vals,
                                              std::unique_ptr<cudf::groupby_aggregation> agg,
                                              cudf::null_policy np = cudf::null_policy::EXCLUDE)
{
  std::vector<cudf::groupby::aggregation_request> requests(1);
  requests[0].values = vals;
  requests[0].aggregations.push_back(std::move(agg));
  cudf::groupby::groupby gb(cudf::table_view({keys}), np);
  auto result = gb.aggregate(requests, cudf::get_default_stream());
  auto k = to_host<K>(result.first->get_column(0).view());
  auto kv = valid_host(result.first->get_column(0).view());
  auto v = to_host<R>(result.second[0].results[0]->view());
  auto vv = valid_host(result.second[0].results[0]->view());
  std::vector<std::tuple<K, R, int>> rows;
  for (size_t i = 0; i < k.size(); ++i) rows.emplace_back(kv[i] ? k[i] : std::numeric_limits<K>::lowest(), v[i], vv[i]);
  std::sort(rows.begin(), rows.end(), [](auto const& a, auto const& b) { return std::get<0>(a) < std::get<0>(b); });
  return rows;
}

template <typename R>
bool same_value(R a, R b)
{
  if constexpr (std::is_floating_point_v<R>) {
    if (std::isnan(a) && std::isnan(b)) return true;
    return a == b || std::fabs(a - b) <= 4 * std::numeric_limits<R>::epsilon() * std::fabs(b);
  } else {
    return a == b;
  }
}

template <typename K, typename R>
void expect(std::vector<std::tuple<K, R, int>> const& got, std::vector<K> keys, std::vector<R> vals,
            std::vector<int> valid = {})
{
  CHECK(got.size() == keys.size());
  if (got.size() != keys.size()) return;
  for (size_t i = 0; i < keys.size(); ++i) {
    int want_valid = valid.empty() ? 1 : valid[i];
    CHECK(std::get<0>(got[i]) == keys[i]);
    CHECK(std::get<2>(got[i]) == want_valid);
    if (want_valid) CHECK(same_value(std::get<1>(got[i]), vals[i]));
  }
}

using K = int32_t;
const std::vector<K> kKeys{1, 2, 3, 1, 2, 2, 1, 3, 3, 2};
const std::vector<K> kNullKeys{1, 2, 3, 1, 2, 2, 1, 3, 3, 2, 4};
const std::vector<int> kNullKeysValid{1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1};
const std::vector<int> kNullValsValid{0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0};

template <typename V>
cudf::type_id tid();
template <> cudf::type_id tid<int8_t>() { return cudf::type_id::INT8; }
template <> cudf::type_id tid<int16_t>() { return cudf::type_id::INT16; }
template <> cudf::type_id tid<int32_t>() { return cudf::type_id::INT32; }
template <> cudf::type_id tid<int64_t>() { return cudf::type_id::INT64; }
template <> cudf::type_id tid<float>() { return cudf::type_id::FLOAT32; }
template <> cudf::type_id tid<double>() { return cudf::type_id::FLOAT64; }

template <typename V>
std::vector<V> iota_vals(int n)
{
  std::vector<V> v(n);
  for (int i = 0; i < n; ++i) v[i] = static_cast<V>(i);
  return v;
}

// sum_tests.cpp:38-53, 106-127
template <typename V, typename R>
void sum_tests()
{
  wrapper<K> keys(cudf::type_id::INT32, kKeys);
  wrapper<V> vals(tid<V>(), iota_vals<V>(10));
  expect<K, R>(run_single<K, R>(keys.view(), vals.view(), cudf::make_sum_aggregation<cudf::groupby_aggregation>()),
               {1, 2, 3}, {9, 19, 17});
  wrapper<K> nk(cudf::type_id::INT32, kNullKeys, kNullKeysValid);
  auto v11 = iota_vals<V>(11);
  v11[10] = 4;
  wrapper<V> nv(tid<V>(), v11, kNullValsValid);
  expect<K, R>(run_single<K, R>(nk.view(), nv.view(), cudf::make_sum_aggregation<cudf::groupby_aggregation>()),
               {1, 2, 3, 4}, {9, 14, 10, 0}, {1, 1, 1, 0});
}

// mean_tests.cpp:48-66, 113-134
template <typename V>
void mean_tests()
{
  wrapper<K> keys(cudf::type_id::INT32, kKeys);
  wrapper<V> vals(tid<V>(), iota_vals<V>(10));
  expect<K, double>(run_single<K, double>(keys.view(), vals.view(),
                                          cudf::make_mean_aggregation<cudf::groupby_aggregation>()),
                    {1, 2, 3}, {3.0, 19.0 / 4, 17.0 / 3});
  wrapper<K> nk(cudf::type_id::INT32, kNullKeys, kNullKeysValid);
  auto v11 = iota_vals<V>(11);
  v11[10] = 4;
  wrapper<V> nv(tid<V>(), v11, kNullValsValid);
  expect<K, double>(run_single<K, double>(nk.view(), nv.view(),
                                          cudf::make_mean_aggregation<cudf::groupby_aggregation>()),
                    {1, 2, 3, 4}, {4.5, 14.0 / 3, 5.0, 0.0}, {1, 1, 1, 0});
}

// min_tests.cpp:37-131, max_tests.cpp:37-129, count_tests.cpp:36-140
template <typename V>
void minmax_count_tests()
{
  wrapper<K> keys(cudf::type_id::INT32, kKeys);
  wrapper<V> vals(tid<V>(), iota_vals<V>(10));
  expect<K, V>(run_single<K, V>(keys.view(), vals.view(), cudf::make_min_aggregation<cudf::groupby_aggregation>()),
               {1, 2, 3}, {0, 1, 2});
  expect<K, V>(run_single<K, V>(keys.view(), vals.view(), cudf::make_max_aggregation<cudf::groupby_aggregation>()),
               {1, 2, 3}, {6, 9, 8});
  expect<K, int32_t>(run_single<K, int32_t>(keys.view(), vals.view(),
                                            cudf::make_count_aggregation<cudf::groupby_aggregation>()),
                     {1, 2, 3}, {3, 4, 3});
  wrapper<K> nk(cudf::type_id::INT32, kNullKeys, kNullKeysValid);
  auto v11 = iota_vals<V>(11);
  v11[10] = 4;
  wrapper<V> nv(tid<V>(), v11, kNullValsValid);
  expect<K, V>(run_single<K, V>(nk.view(), nv.view(), cudf::make_min_aggregation<cudf::groupby_aggregation>()),
               {1, 2, 3, 4}, {3, 1, 2, 0}, {1, 1, 1, 0});
  expect<K, int32_t>(run_single<K, int32_t>(nk.view(), nv.view(),
                                            cudf::make_count_aggregation<cudf::groupby_aggregation>()),
                     {1, 2, 3, 4}, {2, 3, 2, 0});
  expect<K, int32_t>(run_single<K, int32_t>(nk.view(), nv.view(),
                                            cudf::make_count_aggregation<cudf::groupby_aggregation>(
                                              cudf::null_policy::INCLUDE)),
                     {1, 2, 3, 4}, {3, 4, 2, 1});
  std::vector<int> max_valid{1, 1, 1, 1, 1, 1, 0, 1, 1, 0, 0};
  wrapper<V> mv(tid<V>(), v11, max_valid);
  expect<K, V>(run_single<K, V>(nk.view(), mv.view(), cudf::make_max_aggregation<cudf::groupby_aggregation>()),
               {1, 2, 3, 4}, {3, 5, 8, 0}, {1, 1, 1, 0});
}

void keys_tests()
{
  // keys_tests.cpp:98-120: null keys grouped together under INCLUDE
  wrapper<K> nk(cudf::type_id::INT32, kNullKeys, kNullKeysValid);
  auto v = iota_vals<int32_t>(11);
  v[10] = 4;
  wrapper<int32_t> vals(cudf::type_id::INT32, v);
  auto got = run_single<K, int64_t>(nk.view(), vals.view(), cudf::make_sum_aggregation<cudf::groupby_aggregation>(),
                                    cudf::null_policy::INCLUDE);
  // the null key sorts first (its key is replaced by lowest())
  expect<K, int64_t>(got, {std::numeric_limits<K>::lowest(), 1, 2, 3, 4}, {7, 9, 19, 10, 4});
  // sum_tests.cpp:149-170: int32 sums widen to int64
  wrapper<K> zk(cudf::type_id::INT32, {0, 0});
  wrapper<int32_t> ov(cudf::type_id::INT32, {std::numeric_limits<int32_t>::min(), std::numeric_limits<int32_t>::min()});
  expect<K, int64_t>(run_single<K, int64_t>(zk.view(), ov.view(), cudf::make_sum_aggregation<cudf::groupby_aggregation>()),
                     {0}, {-4294967296LL});
  // empty input (sum_tests.cpp:55-70)
  wrapper<K> ek(cudf::type_id::INT32, {});
  wrapper<double> ev(cudf::type_id::FLOAT64, {});
  CHECK((run_single<K, double>(ek.view(), ev.view(), cudf::make_sum_aggregation<cudf::groupby_aggregation>()).empty()));
  // all keys null (sum_tests.cpp:72-87)
  wrapper<K> ak(cudf::type_id::INT32, {1, 2, 3}, {0, 0, 0});
  wrapper<double> av(cudf::type_id::FLOAT64, {3, 4, 5});
  CHECK((run_single<K, double>(ak.view(), av.view(), cudf::make_sum_aggregation<cudf::groupby_aggregation>()).empty()));
}

void multi_request_tests()
{
  // the header example of cpp/include/cudf/groupby.hpp:150-175
  wrapper<int32_t> k1(cudf::type_id::INT32, {1, 2, 1, 3, 1});
  wrapper<int64_t> k2(cudf::type_id::INT64, {1, 2, 1, 4, 1});
  wrapper<int32_t> v(cudf::type_id::INT32, {3, 1, 4, 9, 2});
  std::vector<cudf::groupby::aggregation_request> requests(1);
  requests[0].values = v.view();
  requests[0].aggregations.push_back(cudf::make_sum_aggregation<cudf::groupby_aggregation>());
  requests[0].aggregations.push_back(cudf::make_min_aggregation<cudf::groupby_aggregation>());
  cudf::groupby::groupby gb(cudf::table_view({k1.view(), k2.view()}));
  auto res = gb.aggregate(requests);
  auto a = to_host<int32_t>(res.first->get_column(0).view());
  auto b = to_host<int64_t>(res.first->get_column(1).view());
  auto s = to_host<int64_t>(res.second[0].results[0]->view());
  auto m = to_host<int32_t>(res.second[0].results[1]->view());
  std::map<std::pair<int32_t, int64_t>, std::pair<int64_t, int32_t>> got;
  for (size_t i = 0; i < a.size(); ++i) got[{a[i], b[i]}] = {s[i], m[i]};
  CHECK(got.size() == 3);
  CHECK((got[{1, 1}] == std::pair<int64_t, int32_t>{9, 2}));
  CHECK((got[{2, 2}] == std::pair<int64_t, int32_t>{1, 1}));
  CHECK((got[{3, 4}] == std::pair<int64_t, int32_t>{9, 9}));
  // duplicate aggregations and columns (keys_tests.cpp:367-421)
  wrapper<K> keys(cudf::type_id::INT32, kKeys);
  wrapper<int32_t> vals(cudf::type_id::INT32, iota_vals<int32_t>(10));
  std::vector<cudf::groupby::aggregation_request> dup(2);
  dup[0].values = vals.view();
  dup[0].aggregations.push_back(cudf::make_sum_aggregation<cudf::groupby_aggregation>());
  dup[0].aggregations.push_back(cudf::make_sum_aggregation<cudf::groupby_aggregation>());
  dup[1].values = vals.view();
  dup[1].aggregations.push_back(cudf::make_sum_aggregation<cudf::groupby_aggregation>());
  cudf::groupby::groupby gb2(cudf::table_view({keys.view()}));
  auto r2 = gb2.aggregate(dup);
  CHECK(r2.second.size() == 2 && r2.second[0].results.size() == 2 && r2.second[1].results.size() == 1);
}

// argmax_tests.cpp:91-113, argmin_tests.cpp:94-117, var_tests.cpp:103-146,
// std_tests.cpp:103-122 (the second pass over the rows)
template <typename V>
void second_pass_tests()
{
  wrapper<K> keys(cudf::type_id::INT32, kKeys);
  std::vector<V> desc{9, 8, 7, 6, 5, 4, 3, 2, 1, 0};
  wrapper<V> dv(tid<V>(), desc);
  expect<K, int32_t>(run_single<K, int32_t>(keys.view(), dv.view(),
                                            cudf::make_argmax_aggregation<cudf::groupby_aggregation>()),
                     {1, 2, 3}, {0, 1, 2});
  expect<K, int32_t>(run_single<K, int32_t>(keys.view(), dv.view(),
                                            cudf::make_argmin_aggregation<cudf::groupby_aggregation>()),
                     {1, 2, 3}, {6, 9, 8});
  wrapper<K> ak(cudf::type_id::INT32, kNullKeys, {1, 1, 0, 1, 1, 1, 1, 1, 1, 1, 1});
  std::vector<V> d11{9, 8, 7, 6, 5, 4, 3, 2, 1, 0, 4};
  wrapper<V> av(tid<V>(), d11, {0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 0});
  expect<K, int32_t>(run_single<K, int32_t>(ak.view(), av.view(),
                                            cudf::make_argmax_aggregation<cudf::groupby_aggregation>()),
                     {1, 2, 3, 4}, {3, 4, 7, 0}, {1, 1, 1, 0});
  wrapper<K> nk(cudf::type_id::INT32, kNullKeys, kNullKeysValid);
  auto v11 = iota_vals<V>(11);
  v11[10] = 3;
  wrapper<V> nv2(tid<V>(), v11, {0, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1});
  expect<K, double>(run_single<K, double>(nk.view(), nv2.view(),
                                          cudf::make_variance_aggregation<cudf::groupby_aggregation>()),
                    {1, 2, 3, 4}, {4.5, 49. / 3, 18., 0.}, {1, 1, 1, 0});
  expect<K, double>(run_single<K, double>(nk.view(), nv2.view(),
                                          cudf::make_variance_aggregation<cudf::groupby_aggregation>(2)),
                    {1, 2, 3, 4}, {0., 98. / 3, 0., 0.}, {0, 1, 0, 0});
  expect<K, double>(run_single<K, double>(nk.view(), nv2.view(),
                                          cudf::make_std_aggregation<cudf::groupby_aggregation>()),
                    {1, 2, 3, 4}, {3 / std::sqrt(2.), 7 / std::sqrt(3.), 3 * std::sqrt(2.), 0.}, {1, 1, 1, 0});
}

// quantile_tests.cpp:124-152 (two quantiles: a flattened [groups x 2]
// column), median_tests.cpp:35-53, nunique_tests.cpp:173-198 (INCLUDE),
// nth_element_tests.cpp:94-119
template <typename V>
void sort_path_tests()
{
  wrapper<K> keys(cudf::type_id::INT32, kKeys);
  wrapper<V> vals(tid<V>(), iota_vals<V>(10));
  std::vector<cudf::groupby::aggregation_request> r(1);
  r[0].values = vals.view();
  r[0].aggregations.push_back(
    cudf::make_quantile_aggregation<cudf::groupby_aggregation>({0.25, 0.75}, cudf::interpolation::LINEAR));
  r[0].aggregations.push_back(cudf::make_median_aggregation<cudf::groupby_aggregation>());
  r[0].aggregations.push_back(cudf::make_nth_element_aggregation<cudf::groupby_aggregation>(-2));
  cudf::groupby::groupby gb(cudf::table_view({keys.view()}));
  auto [uk, res] = gb.aggregate(r);
  auto k = to_host<K>(uk->get_column(0).view());
  auto q = to_host<double>(res[0].results[0]->view());
  auto m = to_host<double>(res[0].results[1]->view());
  auto nth = to_host<V>(res[0].results[2]->view());
  CHECK(res[0].results[0]->size() == 6 && q.size() == 6);
  const std::map<K, std::vector<double>> want_q{{1, {1.5, 4.5}}, {2, {3.25, 6.}}, {3, {4.5, 7.5}}};
  const std::map<K, double> want_m{{1, 3.}, {2, 4.5}, {3, 7.}};
  const std::map<K, double> want_nth{{1, 3.}, {2, 5.}, {3, 7.}};
  for (size_t g = 0; g < k.size() && q.size() == 6; ++g) {
    CHECK(same_value(q[2 * g], want_q.at(k[g])[0]) && same_value(q[2 * g + 1], want_q.at(k[g])[1]));
    CHECK(same_value(m[g], want_m.at(k[g])));
    CHECK(same_value(static_cast<double>(nth[g]), want_nth.at(k[g])));
  }
  wrapper<K> k14(cudf::type_id::INT32, {1, 2, 3, 3, 1, 2, 2, 1, 3, 3, 2, 4, 4, 2},
                 {1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 1});
  wrapper<V> v14(tid<V>(), std::vector<V>{0, 1, 2, 2, 3, 4, 5, 6, 7, 8, 9, 4, 4, 2},
                 {0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 1, 0, 0, 0});
  expect<K, int32_t>(run_single<K, int32_t>(k14.view(), v14.view(),
                                            cudf::make_nunique_aggregation<cudf::groupby_aggregation>(
                                              cudf::null_policy::INCLUDE)),
                     {1, 2, 3, 4}, {3, 4, 2, 1});
}

// One list of a lists column: (values, validity) of every element.
template <typename V>
std::vector<std::pair<std::vector<V>, std::vector<int>>> lists_host(cudf::column_view const& c)
{
  cudf::lists_column_view lv(c);
  auto off = to_host<int32_t>(lv.offsets());
  auto vals = to_host<V>(lv.child());
  auto valid = valid_host(lv.child());
  std::vector<std::pair<std::vector<V>, std::vector<int>>> out;
  for (int i = 0; i < lv.size(); ++i)
    out.emplace_back(std::vector<V>(vals.begin() + off[i], vals.begin() + off[i + 1]),
                     std::vector<int>(valid.begin() + off[i], valid.begin() + off[i + 1]));
  return out;
}

// collect_list_tests.cpp:48-65 and collect_set_tests.cpp:357-398 (keys
// {4, 1, 2, 4, 3, 3, 3, 3, 2, 1}): lists in row order; sets ascending, nulls last
template <typename V>
void collect_tests()
{
  wrapper<K> keys(cudf::type_id::INT32, {1, 1, 2, 2, 3, 3});
  wrapper<V> vals(tid<V>(), std::vector<V>{1, 2, 3, 4, 5, 6}, {1, 0, 1, 0, 1, 0});
  std::vector<cudf::groupby::aggregation_request> r(1);
  r[0].values = vals.view();
  r[0].aggregations.push_back(cudf::make_collect_list_aggregation<cudf::groupby_aggregation>());
  r[0].aggregations.push_back(
    cudf::make_collect_list_aggregation<cudf::groupby_aggregation>(cudf::null_policy::EXCLUDE));
  cudf::groupby::groupby gb(cudf::table_view({keys.view()}));
  auto [uk, res] = gb.aggregate(r);
  auto k = to_host<K>(uk->get_column(0).view());
  CHECK(res[0].results[0]->type().id() == cudf::type_id::LIST && res[0].results[0]->num_children() == 2);
  CHECK(res[0].results[0]->child(1).type().id() == tid<V>());
  auto inc = lists_host<V>(res[0].results[0]->view());
  auto exc = lists_host<V>(res[0].results[1]->view());
  CHECK(inc.size() == 3 && exc.size() == 3);
  for (size_t g = 0; g < k.size() && inc.size() == 3 && exc.size() == 3; ++g) {
    const V first = static_cast<V>(2 * k[g] - 1);
    CHECK(inc[g].first.size() == 2 && inc[g].first[0] == first && inc[g].second == std::vector<int>({1, 0}));
    CHECK(exc[g].first == std::vector<V>({first}) && exc[g].second == std::vector<int>({1}));
  }

  wrapper<K> k10(cudf::type_id::INT32, {4, 1, 2, 4, 3, 3, 3, 3, 2, 1});
  wrapper<V> v10(tid<V>(), std::vector<V>{40, 10, 20, 40, 0, 0, 0, 0, 21, 0}, {1, 1, 1, 1, 0, 0, 0, 0, 1, 0});
  std::vector<cudf::groupby::aggregation_request> r2(1);
  r2[0].values = v10.view();
  r2[0].aggregations.push_back(cudf::make_collect_set_aggregation<cudf::groupby_aggregation>());
  r2[0].aggregations.push_back(cudf::make_collect_set_aggregation<cudf::groupby_aggregation>(
    cudf::null_policy::INCLUDE, cudf::null_equality::UNEQUAL));
  r2[0].aggregations.push_back(
    cudf::make_collect_set_aggregation<cudf::groupby_aggregation>(cudf::null_policy::EXCLUDE));
  cudf::groupby::groupby gb2(cudf::table_view({k10.view()}));
  auto [uk2, res2] = gb2.aggregate(r2);
  auto kk = to_host<K>(uk2->get_column(0).view());
  using L = std::pair<std::vector<V>, std::vector<int>>;
  const std::map<K, L> want_eq{{1, {{10, 0}, {1, 0}}}, {2, {{20, 21}, {1, 1}}}, {3, {{0}, {0}}}, {4, {{40}, {1}}}};
  const std::map<K, L> want_ne{
    {1, {{10, 0}, {1, 0}}}, {2, {{20, 21}, {1, 1}}}, {3, {{0, 0, 0, 0}, {0, 0, 0, 0}}}, {4, {{40}, {1}}}};
  const std::map<K, L> want_ex{{1, {{10}, {1}}}, {2, {{20, 21}, {1, 1}}}, {3, {{}, {}}}, {4, {{40}, {1}}}};
  const std::map<K, L>* wants[3] = {&want_eq, &want_ne, &want_ex};
  for (int a = 0; a < 3; ++a) {
    auto got = lists_host<V>(res2[0].results[a]->view());
    CHECK(got.size() == 4);
    for (size_t g = 0; g < kk.size() && got.size() == 4; ++g) {
      const L& w = wants[a]->at(kk[g]);
      CHECK(got[g].second == w.second);
      for (size_t e = 0; e < w.first.size() && got[g].first.size() == w.first.size(); ++e)
        if (w.second[e]) CHECK(got[g].first[e] == w.first[e]);
    }
  }
}

void error_tests()
{
  // groupby.cu:208-212: size mismatch; :176-189: invalid type/aggregation
  wrapper<K> keys(cudf::type_id::INT32, {1, 2, 3});
  wrapper<int32_t> short_vals(cudf::type_id::INT32, {1, 2});
  std::vector<cudf::groupby::aggregation_request> r(1);
  r[0].values = short_vals.view();
  r[0].aggregations.push_back(cudf::make_sum_aggregation<cudf::groupby_aggregation>());
  cudf::groupby::groupby gb(cudf::table_view({keys.view()}));
  bool threw = false;
  try {
    gb.aggregate(r);
  } catch (cudf::logic_error const&) {
    threw = true;
  }
  CHECK(threw);
  wrapper<int32_t> ts(cudf::type_id::TIMESTAMP_DAYS, {1, 2, 3});
  std::vector<cudf::groupby::aggregation_request> r2(1);
  r2[0].values = ts.view();
  r2[0].aggregations.push_back(cudf::make_sum_aggregation<cudf::groupby_aggregation>());
  threw = false;
  try {
    gb.aggregate(r2);
  } catch (cudf::logic_error const&) {
    threw = true;
  }
  CHECK(threw);
}

void hashing_tests()
{
  // spark_murmurhash3_x86_32_test.cpp:221-318 (Spark golden values, seed 42)
  wrapper<int32_t> ints(cudf::type_id::INT32, {0, 100, -100, std::numeric_limits<int32_t>::min(),
                                               std::numeric_limits<int32_t>::max()});
  auto h = cudf::hashing::spark_murmurhash3_x86_32(cudf::table_view({ints.view()}), 42);
  CHECK((to_host<int32_t>(h->view()) == std::vector<int32_t>{933211791, 751823303, -1080202046, 723455942, 133916647}));
  wrapper<int64_t> longs(cudf::type_id::INT64, {0, 100, -100, std::numeric_limits<int64_t>::min(),
                                                std::numeric_limits<int64_t>::max()});
  auto h2 = cudf::hashing::spark_murmurhash3_x86_32(cudf::table_view({longs.view()}), 42);
  CHECK((to_host<int32_t>(h2->view()) ==
         std::vector<int32_t>{-1670924195, 1114849490, 904948192, -853646085, -1604625029}));
}

int main()
{
  struct Suite {
    const char* name;
    std::function<void()> fn;
  };
  std::vector<Suite> suites{
    {"sum<int8>", sum_tests<int8_t, int64_t>},   {"sum<int32>", sum_tests<int32_t, int64_t>},
    {"sum<int64>", sum_tests<int64_t, int64_t>}, {"sum<float>", sum_tests<float, float>},
    {"sum<double>", sum_tests<double, double>},  {"mean<int16>", mean_tests<int16_t>},
    {"mean<int64>", mean_tests<int64_t>},        {"mean<float>", mean_tests<float>},
    {"mean<double>", mean_tests<double>},        {"minmax_count<int8>", minmax_count_tests<int8_t>},
    {"minmax_count<int32>", minmax_count_tests<int32_t>}, {"minmax_count<int64>", minmax_count_tests<int64_t>},
    {"minmax_count<float>", minmax_count_tests<float>},   {"minmax_count<double>", minmax_count_tests<double>},
    {"second_pass<int32>", second_pass_tests<int32_t>},   {"second_pass<double>", second_pass_tests<double>},
    {"second_pass<float>", second_pass_tests<float>},
    {"sort_path<int32>", sort_path_tests<int32_t>},      {"sort_path<double>", sort_path_tests<double>},
    {"collect<int32>", collect_tests<int32_t>},          {"collect<int64>", collect_tests<int64_t>},
    {"collect<double>", collect_tests<double>},
    {"keys", keys_tests},                        {"multi_request", multi_request_tests},
    {"errors", error_tests},                     {"hashing", hashing_tests},
  };
  for (auto& s : suites) {
    int before = g_failures;
    std::printf("[ RUN  ] %s\n", s.name);
    std::fflush(stdout);
    try {
      s.fn();
    } catch (std::exception const& e) {
      ++g_failures;
      std::printf("  FAILED with exception: %s\n", e.what());
    }
    std::printf("[ %s ] %s\n", g_failures == before ? " OK " : "FAIL", s.name);
    std::fflush(stdout);
  }
  std::printf("%d checks, %d failures\n", g_checks, g_failures);
  return g_failures ? 1 : 0;
}
