// admission.h -- the served form's admission rules, host-only and free of
// HIP, so that the CPU tests can drive them with a fake allocator.
//
// A queued shot holds `need` bytes while resident (its slot and the pending
// images it leaves when it finishes); the first shot of a request also brings
// the request's own block (`extra`: coefficients, wavelet, tapers, FP64
// stacks, outputs).  The budget is the HBM the server may hold.
#pragma once

namespace rtm {

enum class Admit { kAdmit, kWait, kReject };

// Decision for the next queued shot.  `releasing`: memory the server holds
// will be released without new admissions (finished requests whose results
// are still being copied out).  A shot that exceeds the budget even alone is
// rejected; one that does not fit now waits while something resident or
// releasing will free memory, and is rejected when nothing will.
inline Admit admit_rule(double budget, double used, int active, bool releasing, double need, double extra)
{
  if (need + extra > budget) return Admit::kReject;
  if (used + need + extra <= budget) return Admit::kAdmit;
  return (active > 0 || releasing) ? Admit::kWait : Admit::kReject;
}

// The device refused an allocation that the budget allowed (the budget is
// above what the device holds): wait while resident or releasing shots will
// return memory, otherwise the request cannot run.
inline Admit alloc_failed_rule(int active, bool releasing)
{
  return (active > 0 || releasing) ? Admit::kWait : Admit::kReject;
}

}  // namespace rtm
