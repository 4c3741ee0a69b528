// Tuple — 16-byte input record; same layout as the reference's Tuple
// (include/sjb/utils/Tuple.hpp:6-11) and as sjb_tuple_t of the C ABI.
#pragma once
#include <cstddef>
#include <cstdint>

#include "sjb.h"

class Tuple {
public:
    uint32_t key;
    uint32_t val;
    uint64_t ts;
};
static_assert(sizeof(Tuple) == sizeof(sjb_tuple_t), "Tuple layout");
static_assert(offsetof(Tuple, key) == offsetof(sjb_tuple_t, key), "Tuple layout");
static_assert(offsetof(Tuple, val) == offsetof(sjb_tuple_t, val), "Tuple layout");
static_assert(offsetof(Tuple, ts) == offsetof(sjb_tuple_t, ts), "Tuple layout");

inline const sjb_tuple_t *as_abi(const Tuple *t) { return reinterpret_cast<const sjb_tuple_t *>(t); }
