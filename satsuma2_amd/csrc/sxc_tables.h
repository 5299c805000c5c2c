// sxc_tables.h -- host construction of the kernels' constant tables (sxc_tables.cpp).
#pragma once
#include "sxc_types.h"

namespace sxc {

void build_tables(sxc_tables* t);
void build_osc(float* osc);  // 24576 floats: oscillator {cos, sin} for FFT levels 13 and 14
int ent_lut_len(int nb);     // doubles in the entropy table for N = 2^nb
void build_ent_lut(int nb, double* out);
int class_of(int byte);
bool class_has_thirds(int c);
// z above which a run's probability is certainly below min_prob (see sxc_scan.h tail).
double keep_z_bound(double min_prob, double target_size);
// ProbTable thresholds for a target size: SXC_PT_ROWS x SXC_PT_LEN doubles (row 0 unused).
void build_prob_table(double target_size, double* t);

}  // namespace sxc
