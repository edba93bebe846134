/* TEST INFRASTRUCTURE ONLY: the native-table query restatement
 * (native_oracle.c) as called by the EM step restatement (em_oracle.c). */
#ifndef C8_ORACLE_NATIVE_INTERNAL_H
#define C8_ORACLE_NATIVE_INTERNAL_H

#include <stdint.h>

#include "c8mi/c8mi_tables.h"

c8mi_native_result oracle_nq_total_rate(const c8mi_native_table_set* v, int32_t pid, double e);
c8mi_native_result oracle_nq_selection_total_rate(const c8mi_native_table_set* v, int32_t pid,
                                                  double e);
c8mi_native_result oracle_nq_rate_for_column(const c8mi_native_table_set* v,
                                             const c8mi_native_dndx_column* c, double e);
int oracle_nq_rate_below_threshold(const c8mi_native_dndx_column* c, double e);
c8mi_native_result oracle_nq_loss_fraction(const c8mi_native_table_set* v, int32_t pid,
                                           int32_t process, uint64_t component, double e,
                                           double q, int photon_split);
c8mi_native_result oracle_nq_range(const c8mi_native_table_set* v, int32_t pid, double e);
c8mi_native_result oracle_nq_energy_after_loss(const c8mi_native_table_set* v, int32_t pid,
                                               double e0, double grammage);
const c8mi_native_utility_column* oracle_nq_find_utility(const c8mi_native_table_set* v,
                                                         int32_t pid);
void oracle_nq_dndx_range(const c8mi_native_table_set* v, int32_t pid, uint32_t* begin,
                          uint32_t* end);
const c8mi_native_dndx_column* oracle_nq_selection_dndx(const c8mi_native_table_set* v,
                                                        uint32_t position);

#endif
