"""The MATLAB gateway (mex/bsamgpu_mex.c) compiles against the library header.

MATLAB is absent, so the gateway is compiled with gcc against declarations of
the C Matrix API functions it calls, written here from their documented
signatures (no MATLAB header or library is used): every call must match them
and every library entry point it names must be declared in include/bsamgpu.h
with the arguments it passes.  The entry points themselves are exercised on
the GPU through the Python layer and the C driver.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GATEWAY = os.path.join(ROOT, "bsam-in-matlab_amd", "mex", "bsamgpu_mex.c")

MATRIX_API = r"""
#include <stddef.h>
#include <stdbool.h>
#include <stdint.h>
typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef bool mxLogical;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
typedef enum { mxUNKNOWN_CLASS, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxVOID_CLASS,
               mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS,
               mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS } mxClassID;
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
int mexAtExit(void (*fn)(void));
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
void mexWarnMsgIdAndTxt(const char* id, const char* fmt, ...);
void mexMakeArrayPersistent(mxArray* a);
void* mxCalloc(size_t n, size_t size);
void* mxRealloc(void* p, size_t size);
void mxFree(void* p);
mxArray* mxCreateCellMatrix(size_t m, size_t n);
mxArray* mxCreateDoubleMatrix(size_t m, size_t n, mxComplexity c);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateLogicalMatrix(size_t m, size_t n);
mxArray* mxCreateNumericMatrix(size_t m, size_t n, mxClassID id, mxComplexity c);
mxArray* mxCreateStructMatrix(size_t m, size_t n, int nfields, const char** names);
mxArray* mxDuplicateArray(const mxArray* a);
void mxDestroyArray(mxArray* a);
mxArray* mxGetCell(const mxArray* a, mwIndex i);
void mxSetCell(mxArray* a, mwIndex i, mxArray* v);
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name);
void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* v);
void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
mxLogical* mxGetLogicals(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
double mxGetScalar(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize len);
bool mxIsChar(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsEmpty(const mxArray* a);
bool mxIsLogical(const mxArray* a);
bool mxIsStruct(const mxArray* a);
bool mxIsUint64(const mxArray* a);
"""


def test_gateway_compiles_against_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    (tmp_path / "mex.h").write_text(MATRIX_API)
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter", "-fsyntax-only",
                        "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"), GATEWAY],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


def test_gateway_maps_every_scheme():
    src = open(GATEWAY).read()
    for name, code in (("be", "BSAM_BE"), ("bdf2", "BSAM_BDF2"), ("cn", "BSAM_CN")):
        assert f'"{name}") ? {code}' in src
    for cmd in ("parabolic_rhs", "parabolic_advance", "parabolic_step", "tag", "cluster", "cluster_per_parent"):
        assert f'"{cmd}"' in src
