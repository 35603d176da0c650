// (T) triples on the device (CcsdPerturbativeTriples.cxx:119-248).
#pragma once

#include <array>
#include <vector>

#include "engine.hpp"

namespace sisi {

// E_(T) over the given sorted blocks (i0 <= i1 <= i2); all blocks when empty.
// In complex mode this is CcsdPerturbativeTriplesComplex
// (CcsdPerturbativeTriplesComplex.cxx:123-279, planar complex amplitudes)
// and the real part of its complex energy is returned (cxx:73-84).
double triples_energy(Engine &e, const double *T1dev, const double *T2dev, int precision,
                      const std::vector<std::array<int, 3>> &blocks);

}  // namespace sisi
