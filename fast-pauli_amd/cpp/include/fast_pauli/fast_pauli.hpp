// fast_pauli.hpp — umbrella header of the MI355X host API.
//
// The same entry points as the reference's fast_pauli.hpp
// (fast_pauli/cpp/include/fast_pauli.hpp:18-23): Pauli, PauliString,
// PauliOp<T, H>, SummedPauliOp<T>, the string enumeration helpers and
// get_sparse_repr.  The compute methods run on the GPU through the C ABI of
// libfastpauli_hip.so (link with -lfastpauli_hip).
#pragma once

#include "device.hpp"
#include "helpers.hpp"
#include "mdspan.hpp"
#include "pauli.hpp"
#include "pauli_op.hpp"
#include "pauli_string.hpp"
#include "summed_pauli_op.hpp"
