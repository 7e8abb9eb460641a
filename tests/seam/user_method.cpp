// A phasing method written the way a maintainer writes one against the reference's seam
// (IPhasingMethod.h:53-69; ePIE.h:55-70 for the override list): it must compile against the
// engine's driver header unchanged.  Compiled (not run) by tests/test_seam.py.
#include "phaser.h"

using namespace ptycho;

class CountingMethod : public IPhasingMethod {
 public:
  unsigned calls = 0;
  void initMem(IPtychoScanMesh* scanMesh, uint2 probeSize) override {
    (void)scanMesh;
    (void)probeSize;
  }
  real_t iteration(Diffractions* diffs, Probe* probe, Sample* object, IPtychoScanMesh* scanMesh,
                   std::vector<std::vector<real_t> >& fourierErrors, const std::vector<float2>& scanPositions,
                   bool phaseConstraint, bool updateProbe, bool updateProbeModes, unsigned int iter,
                   bool RMS = false) override {
    (void)probe, (void)object, (void)scanMesh, (void)scanPositions, (void)updateProbeModes, (void)RMS;
    ++calls;
    if (fourierErrors.size() < iter) fourierErrors.resize(iter);
    uint32_t flags = (phaseConstraint ? PTYK_FLAG_PHASE_CONSTRAINT : 0u) | (updateProbe ? PTYK_FLAG_UPDATE_PROBE : 0u);
    status_ = ptyk_dm_iteration(diffs->ctx, flags);
    return 0;
  }
  void endPhasing() override {}
};

int add_counting_method(IPhaser& ph, unsigned iterations) {
  ph.methods.push_back(CXPhasing{"counting", std::unique_ptr<IPhasingMethod>(new CountingMethod()), iterations});
  return (int)ph.methods.size();
}
