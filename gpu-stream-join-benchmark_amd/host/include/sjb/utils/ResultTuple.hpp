// ResultTuple — 32-byte join result; same layout as the reference's
// ResultTuple (include/sjb/utils/ResultTuple.hpp:5-13) and sjb_result_tuple_t.
#pragma once
#include <cstddef>
#include <cstdint>

#include "sjb.h"

class ResultTuple {
public:
    uint32_t lkey;
    uint32_t rkey;
    uint32_t lVal;
    uint32_t rVal;
    uint64_t lTs;
    uint64_t rTs;
};
static_assert(sizeof(ResultTuple) == sizeof(sjb_result_tuple_t), "ResultTuple layout");
static_assert(offsetof(ResultTuple, rVal) == offsetof(sjb_result_tuple_t, rval), "ResultTuple layout");
static_assert(offsetof(ResultTuple, rTs) == offsetof(sjb_result_tuple_t, rts), "ResultTuple layout");
