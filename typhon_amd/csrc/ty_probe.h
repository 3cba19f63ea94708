// Volume probes (ty_probe.hip): FCT expressions compiled on the host to a postfix program, run per
// cell on the device and reduced to VOL_AVERAGE / VOL_MIN / VOL_MAX.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/typhon_gpu.h"

namespace ty {

// probe types, the reference's codes (CFDTOOLS/Mesh/DEFPROBE.f90:26-28)
constexpr int PROBE_AVG = 21, PROBE_MIN = 22, PROBE_MAX = 23;

constexpr int PROBE_MAXOP = 120;  // program length
constexpr int PROBE_MAXCST = 24;  // constants
constexpr int PROBE_STACK = 16;   // evaluation stack depth

// program words: op 9..13 are the binary + - * / ^ (popping b, then a, pushing a op b); PO_CST
// pushes cst[arg]; PO_VAR pushes variable arg (rho, ps, v, v_x, v_y, v_z, x, y, z, t); PO_FN
// applies unary function arg (0 opposite, 1.. the FCT functions inv .. ramp) to the top
constexpr int PO_CST = 20, PO_VAR = 21, PO_FN = 22;

struct ProbeProg {
    int type;
    int nop, ncst;
    unsigned char op[PROBE_MAXOP];
    unsigned char arg[PROBE_MAXOP];
    double cst[PROBE_MAXCST];
};

struct ProbeData {
    int n, dim;
    const double* u0;   // [nv][n] FP64 conservative state
    const double* cc;   // [3][n] cell centres
    const double* vol;  // [n]
    double g1;          // gamma - 1
    double t;           // current time
};

// parse expr with the FCT grammar; 0 on success, otherwise 1 with the reason in *err
int probe_compile(const char* expr, int type, ProbeProg* prog, std::string* err);
// block partials needed by launch_probe: 2 * probe_blocks(n) doubles
int probe_blocks(int n);
// *out (device) = the probe's value
hipError_t launch_probe(const ProbeProg& p, const ProbeData& d, double* part, double* out, hipStream_t s);

// limits of the allocation-free evaluation's fixed device workspace (2 MiB per GPU)
constexpr int PROBE_SCRATCH_MAX_CELLS = TY_PROBE_NOALLOC_MAX_CELLS;
constexpr int PROBE_SCRATCH_MAX_PROBES = TY_PROBE_NOALLOC_MAX_PROBES;
constexpr size_t PROBE_SCRATCH_DOUBLES = 2 * (PROBE_SCRATCH_MAX_CELLS / 256) + PROBE_SCRATCH_MAX_PROBES;
// device address of the current device's workspace (looked up once per device; the device must
// be current): partials at [0, 2 * probe_blocks(n)), results after 2 * (PROBE_SCRATCH_MAX_CELLS / 256)
hipError_t probe_scratch_setup(int device, double** ptr);

}  // namespace ty
