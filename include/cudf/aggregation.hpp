// Aggregation objects and factories used by cudf::groupby (subset of
// cpp/include/cudf/aggregation.hpp:80-306 and 505-528 in cudf 24.04: same class names,
// Kind numbering and factory signatures).
#pragma once

#include <cudf/types.hpp>

#include <memory>

namespace cudf {

class aggregation {
 public:
  // cpp/include/cudf/aggregation.hpp:85-122
  enum Kind {
    SUM,
    PRODUCT,
    MIN,
    MAX,
    COUNT_VALID,
    COUNT_ALL,
    ANY,
    ALL,
    SUM_OF_SQUARES,
    MEAN,
    M2,
    VARIANCE,
    STD,
    MEDIAN,
    QUANTILE,
    ARGMAX,
    ARGMIN,
    NUNIQUE,
    NTH_ELEMENT,
    ROW_NUMBER,
    RANK,
    COLLECT_LIST,
    COLLECT_SET,
    LEAD,
    LAG,
    PTX,
    CUDA,
    MERGE_LISTS,
    MERGE_SETS,
    MERGE_M2,
    COVARIANCE,
    CORRELATION,
    TDIGEST,
    MERGE_TDIGEST,
    HISTOGRAM,
    MERGE_HISTOGRAM
  };

  aggregation() = delete;
  aggregation(aggregation::Kind a) : kind{a} {}
  Kind kind;
  virtual ~aggregation() = default;
  [[nodiscard]] virtual bool is_equal(aggregation const& other) const { return kind == other.kind; }
  [[nodiscard]] virtual std::unique_ptr<aggregation> clone() const = 0;
};

class rolling_aggregation : public virtual aggregation {
 public:
  ~rolling_aggregation() override = default;

 protected:
  rolling_aggregation() {}
};

class groupby_aggregation : public virtual aggregation {
 public:
  ~groupby_aggregation() override = default;

 protected:
  groupby_aggregation() {}
};

class groupby_scan_aggregation : public virtual aggregation {
 public:
  ~groupby_scan_aggregation() override = default;

 protected:
  groupby_scan_aggregation() {}
};

class reduce_aggregation : public virtual aggregation {
 public:
  ~reduce_aggregation() override = default;

 protected:
  reduce_aggregation() {}
};

template <typename Base = aggregation>
std::unique_ptr<Base> make_sum_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_product_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_min_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_max_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_count_aggregation(null_policy null_handling = null_policy::EXCLUDE);
template <typename Base = aggregation>
std::unique_ptr<Base> make_sum_of_squares_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_mean_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_m2_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_median_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_quantile_aggregation(std::vector<double> const& quantiles,
                                                interpolation interp = interpolation::LINEAR);
template <typename Base = aggregation>
std::unique_ptr<Base> make_variance_aggregation(size_type ddof = 1);
template <typename Base = aggregation>
std::unique_ptr<Base> make_std_aggregation(size_type ddof = 1);
template <typename Base = aggregation>
std::unique_ptr<Base> make_argmax_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_argmin_aggregation();
template <typename Base = aggregation>
std::unique_ptr<Base> make_nunique_aggregation(null_policy null_handling = null_policy::EXCLUDE);
template <typename Base = aggregation>
std::unique_ptr<Base> make_nth_element_aggregation(size_type n, null_policy null_handling = null_policy::INCLUDE);
template <typename Base = aggregation>
std::unique_ptr<Base> make_collect_list_aggregation(null_policy null_handling = null_policy::INCLUDE);
template <typename Base = aggregation>
std::unique_ptr<Base> make_collect_set_aggregation(null_policy null_handling = null_policy::INCLUDE,
                                                   null_equality nulls_equal = null_equality::EQUAL,
                                                   nan_equality nans_equal = nan_equality::ALL_EQUAL);

}  // namespace cudf
