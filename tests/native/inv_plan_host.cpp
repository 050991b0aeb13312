// Host build of the inverse side-task plan of botorch_amd/csrc/inv_plan.h, for
// tests that replay a plan on the CPU.  Test infrastructure only.
#include <cstring>

#include "../../botorch_amd/csrc/inv_plan.h"

// Number of tasks and of steps (NB + 2) of the plan; -1 if it missed a deadline.
extern "C" int ivp_size(int nb, int ncu, int* nsteps) {
  const bta::invplan::Plan P = bta::invplan::build(nb, ncu);
  if (P.off.empty()) return -1;
  *nsteps = (int)P.units.size();
  return (int)P.task.size();
}

// tasks: 4 ints per task; off: nsteps + 1; units, npart: nsteps
extern "C" void ivp_fill(int nb, int ncu, int* tasks, int* off, int* units, int* npart) {
  const bta::invplan::Plan P = bta::invplan::build(nb, ncu);
  if (!P.task.empty()) memcpy(tasks, P.task.data(), P.task.size() * sizeof(P.task[0]));
  memcpy(off, P.off.data(), P.off.size() * sizeof(int));
  memcpy(units, P.units.data(), P.units.size() * sizeof(int));
  memcpy(npart, P.npart.data(), P.npart.size() * sizeof(int));
}

extern "C" int ivp_bulk_tiles(int nb, int j) { return bta::invplan::bulk_tiles(nb, j); }
