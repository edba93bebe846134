/*
 * c8mi native physics tables — the format-v7 table records of CORSIKA8-kokkos's
 * PROPOSAL-native path, layout-compatible with the reference's
 * (corsika/kokkos/em/tables/ProposalNativeData.hpp:99-332), plus the pointer/count
 * set that replaces ProposalNativeDeviceView (:262-332).
 *
 * A table set exported by the reference (ProposalNativeTableSet,
 * corsika/kokkos/em/tables/ProposalNativeTable.hpp:21-56) maps onto
 * c8mi_native_table_set field for field; c8mi copies it to device memory at
 * initialisation.
 */
#ifndef C8MI_TABLES_H
#define C8MI_TABLES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { C8MI_NATIVE_TABLE_FORMAT_VERSION = 7 };
#define C8MI_NO_POLYNOMIAL_COEFFICIENTS UINT64_MAX

/* NativeAxisType */
enum { C8MI_AXIS_LINEAR = 0, C8MI_AXIS_EXPONENTIAL = 1, C8MI_AXIS_EXPONENTIAL_MINUS_ONE = 2 };
/* NativeLossTransform */
enum { C8MI_LOSS_RELATIVE_LOG = 0, C8MI_LOSS_RELATIVE_LOG_POW15 = 1, C8MI_LOSS_ONE_MINUS_LOG = 2 };
/* NativeKinematicModel */
enum {
  C8MI_KIN_UNSUPPORTED = 0,
  C8MI_KIN_ONLY_STOCHASTIC = 1,
  C8MI_KIN_PHOTON_PAIR_FULL_LOSS = 2,
  C8MI_KIN_COMPTON = 3,
  C8MI_KIN_BREMS_ELECTRON_SCREENING = 4,
  C8MI_KIN_ELECTRON_PAIR = 5,
  C8MI_KIN_IONIZATION_MOLLER = 6,
  C8MI_KIN_IONIZATION_BHABHA = 7,
  C8MI_KIN_IONIZATION_BETHE_BLOCH_ROSSI = 8,
  C8MI_KIN_BREMS_KKP = 9,
  C8MI_KIN_PHOTONUCLEAR_ALLM97 = 10
};
/* NativeRateModel */
enum {
  C8MI_RATE_BICUBIC = 0,
  C8MI_RATE_ANNIHILATION_HEITLER = 1,
  C8MI_RATE_PHOTOEFFECT_SAUTER = 2,
  C8MI_RATE_PHOTOPRODUCTION_HECK = 3
};
/* NativeQueryStatus */
enum {
  C8MI_QUERY_SUCCESS = 0,
  C8MI_QUERY_INVALID_VIEW = 1,
  C8MI_QUERY_NON_FINITE_INPUT = 2,
  C8MI_QUERY_PARTICLE_NOT_FOUND = 3,
  C8MI_QUERY_COLUMN_NOT_FOUND = 4,
  C8MI_QUERY_ENERGY_OUT_OF_RANGE = 5,
  C8MI_QUERY_QUANTILE_OUT_OF_RANGE = 6,
  C8MI_QUERY_UNSUPPORTED_KINEMATICS = 7,
  C8MI_QUERY_ROOT_NOT_CONVERGED = 8
};

/* PROPOSAL process identities (corsika/kokkos/em/ProcessId.hpp:15-30). */
enum {
  C8MI_PROCESS_BREMS = 1000000002,
  C8MI_PROCESS_IONIZATION = 1000000003,
  C8MI_PROCESS_ELECTRON_PAIR = 1000000004,
  C8MI_PROCESS_PHOTONUCLEAR = 1000000005,
  C8MI_PROCESS_MUON_PAIR = 1000000006,
  C8MI_PROCESS_COMPTON = 1000000010,
  C8MI_PROCESS_DECAY = 1000000011,
  C8MI_PROCESS_ANNIHILATION = 1000000012,
  C8MI_PROCESS_PHOTON_PAIR = 1000000013,
  C8MI_PROCESS_PHOTOPRODUCTION = 1000000014,
  C8MI_PROCESS_PHOTON_MUON_PAIR = 1000000015,
  C8MI_PROCESS_PHOTOELECTRIC = 1000000016
};

typedef struct c8mi_native_result { /* NativeQueryResult */
  uint32_t status;
  uint32_t newton_iterations;
  uint32_t bisection_iterations;
  uint32_t pad_;
  double value;
} c8mi_native_result;

typedef struct c8mi_native_axis { /* NativeAxisDescriptor */
  uint32_t type;
  uint32_t nodes;
  double low;
  double high;
  double step;
} c8mi_native_axis;

typedef struct c8mi_native_bicubic { /* NativeBicubicDescriptor */
  c8mi_native_axis energy_axis;
  c8mi_native_axis loss_axis;
  uint32_t rows;
  uint32_t columns;
  uint64_t coefficient_offset;
  uint64_t polynomial_coefficient_offset; /* C8MI_NO_POLYNOMIAL_COEFFICIENTS if none */
} c8mi_native_bicubic;

typedef struct c8mi_native_cubic { /* NativeCubicDescriptor */
  c8mi_native_axis axis;
  c8mi_native_axis value_transform;
  uint64_t coefficient_offset;
  uint32_t coefficient_count;
  uint32_t has_value_transform;
} c8mi_native_cubic;

typedef struct c8mi_native_dndx_column { /* NativeDndxColumn */
  int32_t pdg_id;
  int32_t process_id;
  uint64_t medium_hash;
  uint64_t component_hash;
  uint64_t cross_section_hash;
  uint64_t table_hash;
  uint32_t loss_transform;
  uint32_t kinematic_model;
  uint32_t rate_model;
  uint32_t pad_;
  c8mi_native_bicubic spline;
  double component_weight;
  double particle_mass_mev;
  double component_nuclear_charge;
  double component_atomic_mass;
  double component_average_nucleon_mass_mev;
  double direct_rate_factor;
  double lower_energy_limit_mev;
  double absolute_cut_mev;
  double relative_cut;
  double medium_mean_excitation_energy_mev;
  uint32_t selection_ordinal;
  uint32_t reserved;
} c8mi_native_dndx_column;

typedef struct c8mi_native_total_rate_column { /* NativeTotalRateColumn */
  int32_t pdg_id;
  uint32_t reserved;
  uint64_t medium_hash;
  uint64_t table_hash;
  c8mi_native_cubic spline;
  double lower_energy_limit_mev;
} c8mi_native_total_rate_column;

typedef struct c8mi_native_dedx_column { /* NativeDedxColumn */
  int32_t pdg_id;
  int32_t process_id;
  uint64_t medium_hash;
  uint64_t cross_section_hash;
  uint64_t component_index;
  uint64_t table_hash;
  c8mi_native_cubic spline;
  double component_weight;
  double lower_energy_limit_mev;
  uint32_t accumulation_ordinal;
  uint32_t reserved;
} c8mi_native_dedx_column;

typedef struct c8mi_native_utility_column { /* NativeUtilityColumn */
  int32_t pdg_id;
  uint32_t pad0_;
  uint64_t medium_hash;
  uint64_t table_hash;
  c8mi_native_cubic spline;
  double particle_mass_mev;
  double lower_energy_limit_mev;
  uint32_t reverse;
  uint32_t pad1_;
} c8mi_native_utility_column;

typedef struct c8mi_native_constants { /* NativePhysicalConstants */
  double electron_mass_mev;
  double charged_pion_mass_mev;
  double fine_structure_constant;
  double classical_electron_radius_cm;
  double pi;
  double sqrt_e;
  double muon_mass_mev;
} c8mi_native_constants;

/* Host (or device) arrays of one validated format-v7 export; the pointer/count
 * layout of ProposalNativeDeviceView. */
typedef struct c8mi_native_table_set {
  const c8mi_native_dndx_column* dndx_columns;
  uint32_t dndx_column_count;
  uint32_t total_rate_column_count;
  const c8mi_native_total_rate_column* total_rate_columns;
  const uint32_t* selection_column_indices; /* dndx_column_count entries */
  const c8mi_native_dedx_column* dedx_columns;
  const uint32_t* dedx_accumulation_indices; /* dedx_column_count entries */
  uint32_t dedx_column_count;
  uint32_t utility_column_count;
  const c8mi_native_utility_column* utility_columns;
  const double* bicubic_values;
  const double* bicubic_derivative_energy;
  const double* bicubic_derivative_loss;
  const double* bicubic_mixed_derivative;
  uint64_t bicubic_coefficient_count;
  const double* bicubic_polynomial_coefficients; /* may be NULL */
  uint64_t bicubic_polynomial_coefficient_count;
  const double* cubic_values;
  const double* cubic_node_derivatives;
  uint64_t cubic_coefficient_count;
  c8mi_native_constants constants;
} c8mi_native_table_set;

/* Batched table queries (the reference's execute_table_query over a batch,
 * corsika/detail/kokkos/em/tables/TableQuery.inl:12-43, restricted to the
 * PROPOSAL-native kinds). */
enum {
  C8MI_NATIVE_TOTAL_RATE = 0,           /* query_proposal_native_total_rate */
  C8MI_NATIVE_SELECTION_TOTAL_RATE = 1, /* query_proposal_native_selection_total_rate */
  C8MI_NATIVE_RATE = 2,                 /* query_proposal_native_rate */
  C8MI_NATIVE_DEDX = 3,                 /* query_proposal_native_dedx */
  C8MI_NATIVE_LOSS_FRACTION = 4,        /* query_proposal_native_loss_fraction */
  C8MI_NATIVE_PHOTON_SPLIT = 5,         /* loss_fraction with photon_split = true */
  C8MI_NATIVE_RANGE = 6,                /* query_proposal_native_range */
  C8MI_NATIVE_ENERGY_AFTER_LOSS = 7,    /* ..._energy_after_continuous_loss */
  C8MI_NATIVE_ENERGY = 8,               /* query_proposal_native_energy */
  C8MI_NATIVE_CUMULATIVE_RATE = 9       /* ..._cumulative_rate(column of (pid,process,component)) */
};

typedef struct c8mi_native_query {
  uint32_t kind;
  int32_t pdg_id;
  int32_t process_id;
  uint32_t pad_;
  uint64_t component_hash;
  double energy_mev;
  double argument; /* quantile, grammage or range, by kind */
} c8mi_native_query;

/* ---- the table set on disk ----
 * The metadata a ProposalNativeTableSet carries beside its arrays
 * (corsika/kokkos/em/tables/ProposalNativeTable.hpp:21-56): format version,
 * dependency versions, stochastic cut and the Interaction identities. */
typedef struct c8mi_native_interaction_identity { /* NativeInteractionIdentity */
  int32_t pdg_id;
  uint32_t reserved;
  uint64_t medium_hash;
  uint64_t interaction_hash;
} c8mi_native_interaction_identity;

typedef struct c8mi_native_table_meta {
  uint32_t format_version;               /* 7 */
  uint32_t pad_;
  char proposal_version[32];             /* NUL-terminated, "7.6.2" for a PROPOSAL export */
  char cubic_interpolation_version[32];  /* "0.1.5" */
  double stochastic_cut_mev;
  const c8mi_native_interaction_identity* interaction_identities;
  uint64_t interaction_identity_count;
} c8mi_native_table_meta;

/* sha256 (corsika/kokkos/utility/Sha256.hpp:21): FIPS 180-4 SHA-256 of a byte
 * string. */
int c8mi_sha256(const uint8_t* data, uint64_t size, uint8_t digest[32]);

/* calculate_proposal_native_hash (src/kokkos/em/tables/ProposalNativeTable.cpp:
 * 570-666): SHA-256 of the canonical little-endian encoding of the metadata
 * and every array, in the reference's order; the table's content hash (the
 * `table_sha256` of the shower report). */
int c8mi_native_table_hash(const c8mi_native_table_set* set, const c8mi_native_table_meta* meta,
                           uint8_t digest[32]);

/* The table file: 8 bytes "C8MINT07", the payload length (u64 little
 * endian), the canonical encoding above as payload, its SHA-256. Written to a
 * temporary file in the target's directory, read back and verified, then
 * renamed over `path`. */
int c8mi_native_table_write(const char* path, const c8mi_native_table_set* set,
                            const c8mi_native_table_meta* meta);

/* Reads and verifies a table file (magic, length, digest, every descriptor
 * within its arrays): an owning handle whose set and metadata views stay
 * valid until it is destroyed; pass the set to c8mi_em_create or
 * c8mi_em_tables_create. A bad file is status 1 (invalid_argument), an
 * unreadable one status 4. */
typedef struct c8mi_native_table_file c8mi_native_table_file;
int c8mi_native_table_read(const char* path, c8mi_native_table_file** out);
int c8mi_native_table_file_view(const c8mi_native_table_file* file, c8mi_native_table_set* set,
                                c8mi_native_table_meta* meta, uint8_t digest[32]);
int c8mi_native_table_file_destroy(c8mi_native_table_file* file);

#ifdef __cplusplus
static_assert(sizeof(c8mi_native_axis) == 32, "NativeAxisDescriptor layout");
static_assert(sizeof(c8mi_native_bicubic) == 88, "NativeBicubicDescriptor layout");
static_assert(sizeof(c8mi_native_cubic) == 80, "NativeCubicDescriptor layout");
#endif

#ifdef __cplusplus
}
#endif

#endif /* C8MI_TABLES_H */
