// Cut-sample vectors of minkasi's CutsCompact (minkasi/tods/cuts.py:201-498):
// the samples a TOD marks as cut are solved for as free parameters, one
// value each, addressed by imap, the flat (det * ndata + sample) index of
// every cut sample (cuts.py:330-347).
//
//  tod2cuts  : vec[i] (+)= dat[imap[i]]         (minkasi.c:260-268)
//  cuts2tod  : dat[imap[i]] (+)= vec[i]         (minkasi.c:271-284)
//
// One lane per cut sample.  Cut samples come in runs (segments), so the
// dat accesses of a wave are mostly contiguous.  Indices outside
// [0, nsamp) are skipped and counted into *nbad (the reference does not
// check them).  cuts2tod requires imap without repeated entries: two lanes
// would race on one sample.  CutsCompact only builds such maps when its
// segments do not overlap (merge_cuts), and its host side checks.
#include "common.h"

namespace {

template <typename T>
__global__ void __launch_bounds__(256) tod2cuts_kernel(T* __restrict__ vec, const T* __restrict__ dat, int64_t nsamp,
                                                       const int64_t* __restrict__ imap, int64_t ncut, int do_add,
                                                       int64_t* nbad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncut; i += stride) {
    const int64_t j = imap[i];
    const bool bad = (uint64_t)j >= (uint64_t)nsamp;
    mk::count_bad(bad, nbad);
    if (bad) continue;
    const T v = dat[j];
    vec[i] = do_add ? vec[i] + v : v;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) cuts2tod_kernel(T* __restrict__ dat, const T* __restrict__ vec, int64_t nsamp,
                                                       const int64_t* __restrict__ imap, int64_t ncut, int do_add,
                                                       int64_t* nbad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncut; i += stride) {
    const int64_t j = imap[i];
    const bool bad = (uint64_t)j >= (uint64_t)nsamp;
    mk::count_bad(bad, nbad);
    if (bad) continue;
    const T v = vec[i];
    dat[j] = do_add ? dat[j] + v : v;
  }
}

template <typename T>
int tod2cuts_impl(T* vec, const T* dat, int64_t nsamp, const int64_t* imap, int64_t ncut, int do_add,
                  int64_t* nbad, void* stream) {
  MK_REQUIRE(ncut >= 0 && nsamp >= 0, MK_ERR_ARG, "negative size");
  if (ncut == 0) return MK_OK;
  MK_REQUIRE(vec && dat && imap, MK_ERR_ARG, "null pointer");
  tod2cuts_kernel<T><<<mk::grid_for(ncut, 256, 256 * 32), 256, 0, mk::as_stream(stream)>>>(vec, dat, nsamp, imap,
                                                                                            ncut, do_add, nbad);
  return mk::check_launch("tod2cuts_kernel");
}

template <typename T>
int cuts2tod_impl(T* dat, const T* vec, int64_t nsamp, const int64_t* imap, int64_t ncut, int do_add,
                  int64_t* nbad, void* stream) {
  MK_REQUIRE(ncut >= 0 && nsamp >= 0, MK_ERR_ARG, "negative size");
  if (ncut == 0) return MK_OK;
  MK_REQUIRE(vec && dat && imap, MK_ERR_ARG, "null pointer");
  cuts2tod_kernel<T><<<mk::grid_for(ncut, 256, 256 * 32), 256, 0, mk::as_stream(stream)>>>(dat, vec, nsamp, imap,
                                                                                            ncut, do_add, nbad);
  return mk::check_launch("cuts2tod_kernel");
}

}  // namespace

extern "C" int mk_tod2cuts_f32(float* vec, const float* dat, int64_t nsamp, const int64_t* imap, int64_t ncut,
                               int do_add, int64_t* dev_nbad, void* stream) {
  return tod2cuts_impl<float>(vec, dat, nsamp, imap, ncut, do_add, dev_nbad, stream);
}
extern "C" int mk_tod2cuts_f64(double* vec, const double* dat, int64_t nsamp, const int64_t* imap, int64_t ncut,
                               int do_add, int64_t* dev_nbad, void* stream) {
  return tod2cuts_impl<double>(vec, dat, nsamp, imap, ncut, do_add, dev_nbad, stream);
}
extern "C" int mk_cuts2tod_f32(float* dat, const float* vec, int64_t nsamp, const int64_t* imap, int64_t ncut,
                               int do_add, int64_t* dev_nbad, void* stream) {
  return cuts2tod_impl<float>(dat, vec, nsamp, imap, ncut, do_add, dev_nbad, stream);
}
extern "C" int mk_cuts2tod_f64(double* dat, const double* vec, int64_t nsamp, const int64_t* imap, int64_t ncut,
                               int do_add, int64_t* dev_nbad, void* stream) {
  return cuts2tod_impl<double>(dat, vec, nsamp, imap, ncut, do_add, dev_nbad, stream);
}
