// functional_types.h -- the problem ("plugin") interface of the GA* search.
//
// Same function-pointer types as the reference's functional_types.h:6-25, so a
// problem written against the reference plugs into gastar/a_star.h unchanged:
//   NeighbourGenerator<T>: writes successor i of `source` into `target`, sets
//                          `target_exists` and the move `cost` (i in [0, K)).
//   Heuristic<T>:          admissible lower bound on the remaining cost.
//   Hash<T>:               declared for source compatibility only.  Neither
//                          the reference's a_star<> nor this one takes a
//                          hash parameter: the reference hard-wires
//                          MurmurHash3 (helpers.h:10-35), this engine hashes
//                          a state's bytes with its own 64-bit mix
//                          (gastar::hash_words, gastar::mix64).
#ifndef GASTAR_FUNCTIONAL_TYPES_H
#define GASTAR_FUNCTIONAL_TYPES_H

#include <cstdint>

template <typename T>
using NeighbourGenerator = void (*)(const T& source, T& target, bool& target_exists, uint32_t& cost,
                                    unsigned char i);

template <typename T>
using Heuristic = uint32_t (*)(const T&);

template <typename T>
using BinOp = T (*)(const T&, const T&);

template <typename T, typename U>
using UnOp = U (*)(const T&);

template <typename T>
using OrdFun = bool (*)(const T&, const T&);

template <typename T>
using EqFun = bool (*)(const T&, const T&);

template <typename T>
using Hash = uint64_t (*)(const T&);

#endif
