// Launch wrappers of the V-cycle kernels (bsam_kernels.hip).
#pragma once
#include "bsam_internal.h"

namespace bsam {

template <typename T> void launch_relax(const LevelDev<T>& L, int colour, T omega, int homog, hipStream_t s);
template <typename T> void launch_fill(const LevelDev<T>& L, const LevelDev<T>* parent, int homog, hipStream_t s);
template <typename T> void launch_ring(const LevelDev<T>& L, const LevelDev<T>& parent, hipStream_t s);
template <typename T> void launch_restrict(const LevelDev<T>& L, const LevelDev<T>& parent, int qc_only, hipStream_t s);
template <typename T> void launch_coarse_load(const LevelDev<T>& L, const LevelDev<T>& parent, int mass_corr, hipStream_t s);
template <typename T> void launch_correct(const LevelDev<T>& L, const LevelDev<T>& parent, int init, hipStream_t s);
template <typename T> void launch_coarsest(const CoarsestDev& C, const LevelDev<T>& L1, int homog, hipStream_t s);
template <typename T, typename TO>
void launch_residual(const LevelDev<T>& L, const LevelDev<T>* fine, int mass_corr, ResidualOut* parts, TO* rout,
                     hipStream_t s);
void launch_reduce(const ResidualOut* parts, int nparts, ResidualOut* out, hipStream_t s);
template <typename T> void launch_mass(const LevelDev<T>& L, ResidualOut* parts, hipStream_t s);
void launch_add_correction(const LevelDev<double>& L, const LevelDev<float>& E, hipStream_t s);
template <typename T> void launch_fill_value(T* a, long long n, T v, hipStream_t s);
template <typename TD, typename TS> void launch_convert(TD* d, const TS* src, long long n, hipStream_t s);
void launch_bdf2_rhs(double* FT, const double* f, const double* Uo, const double* Uoo, double dt, int first,
                     long long n, hipStream_t s);
void launch_history(const LevelDev<double>& L, double* Uo, double* Uoo, int rotate, hipStream_t s);
void launch_invdiag(const LevelDev<double>& L, float* ID32, hipStream_t s);
void launch_add_scalar(double* a, long long n, double sft, float* a32, hipStream_t s);

}  // namespace bsam
