// params.cpp -- the DICpara fields the local matching path reads, as a plain
// struct (host code, no GPU work): defaults, validation, and the search-radius
// rule of IntegerSearchQuadtree.  The values and rules follow the reference's
// MATLAB (cited per function); the interactive prompts have no counterpart.
#include <cstdint>
#include <string>

#include "../../include/aldic.h"

namespace aldic {
int set_error(int code, const std::string& msg);
}

namespace {
bool pow2(int32_t v) { return v > 0 && (v & (v - 1)) == 0; }

// a search distance of n values: 0 is empty ([]), otherwise 1, 2 or 4
// positive values
bool distance_ok(const int32_t* d, int32_t n) {
  if (n == 0) return true;
  if (n != 1 && n != 2 && n != 4) return false;
  for (int32_t k = 0; k < n; k++)
    if (d[k] < 0) return false;
  return true;
}
}  // namespace

// setDICparaDefaults.m:18-56 (the fields the path reads), with the drivers'
// tolerances and ALM constants: StereoMatch_STAQ.m:71 (tol 1e-3),
// TemporalMatch_quadtree_ST1.m:190 (tol 1e-2), :266 (mu), :272 (alpha),
// :298 (beta = 1e-3 mean(winstepsize)^2 mu when DICpara.beta is absent:
// left at 0 here, meaning "derive"), :333 (tol2).
extern "C" int aldic_params_default(aldic_params* p) {
  if (!p) return aldic::set_error(ALDIC_EINVAL, "params_default: NULL");
  *p = aldic_params{};
  p->winsize = 0;       // [] -- the caller sets it
  p->winstepsize = 0;   // []
  p->winsizeMin = 8;
  p->DICIncOrNot = 0;
  p->UseGlobal = 1;
  p->GaussPtOrder = 2;
  p->ClusterNo = 1;
  p->InitFFTSearchMethod = 1;
  p->fixSearchDistanceOrNot = 0;
  p->tol_stereo = 1e-3;
  p->tol_temporal = 1e-2;
  p->mu = 1e-3;
  p->beta = 0.0;
  p->alpha = 0.0;
  p->tol2 = 1e-2;
  return ALDIC_OK;
}

// validateDICpara.m:19-98, the rules on these fields: an even positive
// winsize; winstepsize and winsizeMin powers of 2 with winsizeMin <=
// winstepsize (checked when winstepsize is set); DICIncOrNot 0 or 1;
// InitFFTSearchMethod 0..3; ClusterNo >= 1; UseGlobal 0 or 1; search
// distances empty, 1, 2 or 4 values (:93-98 says positive; 0 is accepted
// here, as IntegerSearchQuadtree uses it).  Unset (0) winsize and
// winstepsize are skipped, as the reference skips empty ones.
extern "C" int aldic_params_validate(const aldic_params* p) {
  if (!p) return aldic::set_error(ALDIC_EINVAL, "params_validate: NULL");
  if (p->winsize != 0 && (p->winsize < 0 || p->winsize % 2 != 0))
    return aldic::set_error(ALDIC_EINVAL, "winsize must be a positive even integer");
  if (p->winstepsize != 0) {
    if (!pow2(p->winstepsize))
      return aldic::set_error(ALDIC_EINVAL, "winstepsize must be a positive power of 2");
    if (!pow2(p->winsizeMin))
      return aldic::set_error(ALDIC_EINVAL, "winsizeMin must be a positive power of 2");
    if (p->winsizeMin > p->winstepsize)
      return aldic::set_error(ALDIC_EINVAL, "winsizeMin must be <= winstepsize");
  }
  if (p->DICIncOrNot != 0 && p->DICIncOrNot != 1)
    return aldic::set_error(ALDIC_EINVAL, "DICIncOrNot must be 0 or 1");
  if (p->InitFFTSearchMethod < 0 || p->InitFFTSearchMethod > 3)
    return aldic::set_error(ALDIC_EINVAL, "InitFFTSearchMethod must be 0-3");
  if (p->ClusterNo < 1) return aldic::set_error(ALDIC_EINVAL, "ClusterNo must be >= 1");
  if (p->UseGlobal != 0 && p->UseGlobal != 1)
    return aldic::set_error(ALDIC_EINVAL, "UseGlobal must be 0 or 1");
  if (!distance_ok(p->StereoSearchDistance, p->nStereoSearchDistance) ||
      !distance_ok(p->TemporalSearchDistance, p->nTemporalSearchDistance) ||
      !distance_ok(p->NewFFTSearchDistance, p->nNewFFTSearchDistance))
    return aldic::set_error(ALDIC_EINVAL,
                            "a search distance must be empty, a scalar, [x y] or [L R U D]");
  return ALDIC_OK;
}

// IntegerSearchQuadtree.m:52-90 and funIntegerSearch.m:12: stereo matching
// (ImgSeqNum == -1) takes StereoSearchDistance, else NewFFTSearchDistance;
// temporal matching (ImgSeqNum >= 2) TemporalSearchDistance, else
// NewFFTSearchDistance; a scalar becomes [s s]; [x y] becomes [x y x y]
// (repmat); [L R U D] stays.  Where the reference would prompt with input()
// (no preset, or ImgSeqNum 0 or 1) this is ALDIC_EINVAL.
extern "C" int aldic_params_search_region(const aldic_params* p, int32_t ImgSeqNum,
                                          int64_t region[4]) {
  if (!p || !region) return aldic::set_error(ALDIC_EINVAL, "params_search_region: NULL");
  const int32_t* d = nullptr;
  int32_t n = 0;
  if (ImgSeqNum == -1) {
    if (p->nStereoSearchDistance > 0) {
      d = p->StereoSearchDistance;
      n = p->nStereoSearchDistance;
    } else if (p->nNewFFTSearchDistance > 0) {
      d = p->NewFFTSearchDistance;
      n = p->nNewFFTSearchDistance;
    }
  } else if (ImgSeqNum >= 2) {
    if (p->nTemporalSearchDistance > 0) {
      d = p->TemporalSearchDistance;
      n = p->nTemporalSearchDistance;
    } else if (p->nNewFFTSearchDistance > 0) {
      d = p->NewFFTSearchDistance;
      n = p->nNewFFTSearchDistance;
    }
  }
  if (!d)
    return aldic::set_error(ALDIC_EINVAL,
                            "search distance not set (the reference would prompt with input())");
  if (!distance_ok(d, n))
    return aldic::set_error(ALDIC_EINVAL, "bad search distance");
  if (n == 1) {
    for (int k = 0; k < 4; k++) region[k] = d[0];
  } else if (n == 2) {
    region[0] = d[0];
    region[1] = d[1];
    region[2] = d[0];
    region[3] = d[1];
  } else {
    for (int k = 0; k < 4; k++) region[k] = d[k];
  }
  return ALDIC_OK;
}
