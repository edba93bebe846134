// Host side of Molière multiple scattering: snapshot validation and the
// interpolation table (construction, reference cache file).
#pragma once

#include "c8mi/c8mi_em.h"

namespace c8mi::em {

// valid_moliere_snapshot (MoliereSnapshot.inl:26-66): finite, positive constants
// and components, finite coefficients.
bool valid_moliere_snapshot(c8mi_moliere_snapshot const& m);
// Every polynomial coefficient and initial-guess correction finite
// (finite_interpolation_table, MoliereInterpolation.cpp:193-203).
bool valid_moliere_table(c8mi_moliere_interpolation const& table);
void build_moliere_table(c8mi_moliere_snapshot const& snapshot, c8mi_moliere_interpolation& out);
// False when the file is missing, incomplete or does not match the snapshot.
bool read_moliere_table(char const* path, c8mi_moliere_snapshot const& snapshot, c8mi_moliere_interpolation& out);
void write_moliere_table(char const* path, c8mi_moliere_snapshot const& snapshot,
                         c8mi_moliere_interpolation const& table);

}  // namespace c8mi::em
