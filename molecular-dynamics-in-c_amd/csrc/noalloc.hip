// noalloc.hip -- mdg_forces_noalloc: mdc's calc_foc (rv.c:12-18) for one box
// whose positions, species, forces and energies are the caller's device
// buffers, with every other device buffer (the box's partial sums, phase
// tables, structure-factor partials, cell lists, k-space plan, pair table,
// descriptor and work lists) in `scratch`, one file-scope device array.
// The call allocates no device memory and creates no stream or event.
//
// `scratch` is a host-side stand-in in host code: its device address is
// taken once per device with hipGetSymbolAddress (with that device
// current) and kept in g_dev[dev].base; only that pointer is ever handed
// to kernels or copies.  Every call on a device shares its scratch, which
// keeps the last call's box: a call with the same parameters and precision
// reuses the uploaded plan, any other call rebuilds it.  Calls that
// overlap on one device race; calls are serialised within the process.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "engine.h"
#include "mdgpu.h"

namespace mdg {
int spec_from_params(const mdg_box_params& p, BoxSpec* s, std::vector<double>* pe,
                     std::vector<double>* fe);
}

using namespace mdg;

// MDG_NOALLOC_SCRATCH_BYTES of device memory per device, held until the
// process ends.  A 4096-ion NaCl box (Ewald at 1e-6, rc = 10 A) needs
// about 18 MiB of it in either precision.
static __device__ __align__(256) unsigned char scratch[MDG_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
constexpr int kCheckThreads = 256;
constexpr int kErrSpecies = 4;   // a species outside 0 .. kinds-1

struct DevState {
  void* base = nullptr;   // this device's copy of scratch (hipGetSymbolAddress)
  Box* box = nullptr;     // the box of the last call, laid out at base
  BoxSet* set = nullptr;  // its work lists, in scratch after the box
  mdg_box_params key{};
  int precision = -1;
};

DevState g_dev[kMaxDevices];
std::mutex g_mu;

bool same_params(const mdg_box_params& a, const mdg_box_params& b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.ax == b.ax && a.ay == b.ay &&
         a.az == b.az && a.alpha == b.alpha && a.hm == b.hm && a.rc == b.rc &&
         a.dt_fs == b.dt_fs && a.real_mode == b.real_mode && a.cells == b.cells &&
         memcmp(&a.pot, &b.pot, sizeof(a.pot)) == 0;
}

// Before any force kernel reads them: every species within 0 .. kinds-1
// (the pair-type lookup and the table index depend on it) and every
// coordinate finite.  Also mdc's self term -a3 sum Z^2 (rcprcl.c:119-125)
// from the device species, summed in a fixed order, into the device
// descriptor that k_box_scalars reads.
__global__ __launch_bounds__(kCheckThreads) void k_noalloc_check(BoxDesc* __restrict__ boxes) {
  BoxDesc& B = boxes[0];
  const int tid = threadIdx.x;
  double z2 = 0.0;
  int bad = 0;
  for (int i = tid; i < B.N; i += kCheckThreads) {
    const int t = B.typ[i];
    if (t < 0 || t >= B.kinds) {
      bad |= kErrSpecies;
    } else {
      z2 += B.Z[t] * B.Z[t];
    }
    if (!(isfinite(B.rx[i]) && isfinite(B.ry[i]) && isfinite(B.rz[i]))) bad |= kErrState;
  }
  __shared__ double s_z[kCheckThreads];
  __shared__ int s_bad[kCheckThreads];
  s_z[tid] = z2;
  s_bad[tid] = bad;
  __syncthreads();
  for (int off = kCheckThreads / 2; off > 0; off >>= 1) {
    if (tid < off) {
      s_z[tid] += s_z[tid + off];
      s_bad[tid] |= s_bad[tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    B.self_pot = -s_z[0] * B.a3;
    if (s_bad[0]) atomicOr(B.err, s_bad[0]);
  }
}

int read_err(const Box* b, hipStream_t s, int* e) {
  MDG_HIP(hipMemcpyAsync(e, b->desc.err, sizeof(int), hipMemcpyDeviceToHost, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

}  // namespace

extern "C" int mdg_forces_noalloc(const mdg_box_params* p, int precision, const double* r,
                                  const int* species, double* f, double* energies,
                                  void* stream) {
  if (!p || !r || !species || !f || !energies) return MDG_E_ARG;
  if (precision != MDG_FP64 && precision != MDG_MIXED) return MDG_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return MDG_E_NODEVICE;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  DevState& d = g_dev[dev];
  if (!d.base) {  // once per device, with that device current
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_last_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol,
                         "hipGetSymbolAddress(scratch)");
      return MDG_E_HIP;
    }
    d.base = ptr;
  }
  if (!d.box || d.precision != precision || !same_params(d.key, *p)) {
    delete d.set;
    d.set = nullptr;
    box_destroy(d.box);  // host bookkeeping only: the box lives in scratch
    d.box = nullptr;
    d.precision = -1;
    BoxSpec spec;
    std::vector<double> pe, fe;
    int rc = spec_from_params(*p, &spec, &pe, &fe);
    if (rc != MDG_OK) return rc;
    Box* box = nullptr;
    std::string why;
    rc = box_create_at(spec, precision, d.base, MDG_NOALLOC_SCRATCH_BYTES, s, &box, &why);
    if (rc != MDG_OK) return rc;  // MDG_E_BUDGET: the box needs more than the scratch
    const size_t used = (box->bytes + 255) & ~(size_t)255;
    if (used >= MDG_NOALLOC_SCRATCH_BYTES) {
      box_destroy(box);
      return MDG_E_BUDGET;
    }
    BoxSet* set = new BoxSet(precision, p->real_mode);
    set->set_arena((char*)d.base + used, MDG_NOALLOC_SCRATCH_BYTES - used);
    rc = set->set_boxes({box});
    if (rc != MDG_OK) {
      delete set;
      box_destroy(box);
      return rc;
    }
    d.box = box;
    d.set = set;
    d.key = *p;
    d.precision = precision;
  }
  Box* box = d.box;
  BoxDesc& D = box->desc;
  const int N = D.N;
  // the caller's buffers in place of the box's own r, f and species
  D.rx = const_cast<double*>(r);
  D.ry = D.rx + N;
  D.rz = D.rx + 2 * (size_t)N;
  D.fx = f;
  D.fy = f + N;
  D.fz = f + 2 * (size_t)N;
  D.typ = const_cast<int*>(species);
  box->desc_dirty = true;
  int rc = d.set->sync_desc(s);
  if (rc != MDG_OK) return rc;
  // the error flags persist in scratch between calls: clear them first
  MDG_HIP(hipMemsetAsync(D.err, 0, sizeof(int), s));
  hipLaunchKernelGGL(k_noalloc_check, dim3(1), dim3(kCheckThreads), 0, s, d.set->device_desc());
  MDG_HIP(hipGetLastError());
  int e = 0;
  rc = read_err(box, s, &e);
  if (rc != MDG_OK) return rc;
  if (e & kErrSpecies) return MDG_E_ARG;
  if (e & kErrState) return MDG_E_STATE;
  d.set->enqueue_step(s, false);
  MDG_HIP(hipGetLastError());
  // pot, pot_real, pot_recip, pot_self, etot (kinetic energy 0), vir x, y, z
  MDG_HIP(hipMemcpyAsync(energies, D.scal, sizeof(double) * 8, hipMemcpyDeviceToDevice, s));
  rc = read_err(box, s, &e);
  if (rc != MDG_OK) return rc;
  if (e & kErrState) return MDG_E_STATE;
  if (e & kErrRange) return MDG_E_RANGE;
  return MDG_OK;
}
