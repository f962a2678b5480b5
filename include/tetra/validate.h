/*
 * validate.h -- host-vs-device relation checks with the reference's names
 * and tolerances (include/validate.h:7-15, src/validate.cpp:40-613): EV
 * accepts a reversed pair, TF and FV compare sorted rows, TE accepts any
 * permutation of a row, TT/FF/EE/VV check that every host entry appears in
 * the device row (VV tolerates duplicates).  Each prints a summary and
 * returns true when the device relation passes.
 *
 * check_exact_* are this library's strict forms: identical sizes, rows and
 * order (VV after sorting rows, as its order within a row is unspecified).
 */
#ifndef TETRA_VALIDATE_H
#define TETRA_VALIDATE_H

#include "datatypes.h"

bool check_host_vs_device_TE(const TE_Data &host_TE, const TE_Data &device_TE);
bool check_host_vs_device_EV(const EV_Data &host_EV, const EV_Data &device_EV);
bool check_host_vs_device_TF(TF_Data &host_TF, TF_Data &device_TF);
bool check_host_vs_device_FV(FV_Data &host_FV, FV_Data &device_FV);
bool check_host_vs_device_TT(const TT_Data &host_TT, const TT_Data &device_TT);
bool check_host_vs_device_FF(const FF_Data &host_FF, const FF_Data &device_FF);
bool check_host_vs_device_EE(const EE_Data &host_EE, const EE_Data &device_EE);
bool check_host_vs_device_VV(const VV_Data &host_VV, const VV_Data &device_VV);

template <typename T>
bool check_exact(const char *name, const T &host, const T &device);
bool check_exact_VV(const VV_Data &host_VV, const VV_Data &device_VV);
bool check_exact_classes(const unsigned int *host, const unsigned int *device,
                         vtkIdType n_points);

#endif
