"""The source-compatible device TimeLoop (include/octlb_gpu_time_loop.h)
compiles against the reference's own headers.

A translation unit written the way a reference case would use the shim --
the reference's constructor arguments for TimeLoop, the boundary handler and
the three inlet profiles, every public member -- is checked with
g++ -std=c++20 -fsyntax-only against the reference tree (full semantic
analysis: overloads, conversions, virtual overrides).  It cannot link: the
reference's OctreeForest needs p4est, which is absent.  Skipped where the
reference sources are not present.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")

USE = r'''
#include "octlb_gpu_time_loop.h"

#include <memory>
#include <vector>

namespace o = octlb;

// Cylinder3dCase-style wiring (cylinder3d_case.h:392-469) with the device
// replacements: unchanged constructor arguments.
double Drive(const o::OctreeForest& forest, const o::FacePairList& faces,
             o::BlockCollection<o::TimeLoopLattice>& blocks, o::GhostSchedule<o::TimeLoopLattice>& ghosts,
             o::LevelCoupler& coupler, int n, double omega) {
  std::vector<o::DomainBcSpec> specs(3);
  specs[0].face = o::FaceDir::kXMin;
  specs[0].type = o::DomainBcType::kInterpolatedVelocity;
  specs[0].inlet_field = std::make_shared<o::gpu::UniformInletProfile>(std::array<double, 3>{0.04, 0.0, 0.0}, 100.0);
  specs[1].face = o::FaceDir::kYMin;
  specs[1].type = o::DomainBcType::kInterpolatedVelocity;
  specs[1].inlet_field = std::make_shared<o::gpu::PoiseuilleInletProfile>(
      o::FaceDir::kYMin, n, n, 0.05, std::array<double, 3>{0.0, 1.0, 0.0}, 0.0);
  specs[2].face = o::FaceDir::kXMax;
  specs[2].type = o::DomainBcType::kInterpolatedPressure;
  specs[2].rho_target = 1.0;
  std::vector<o::DomainBcSpec> more = specs;
  more[1].inlet_field = std::make_shared<o::gpu::ChannelPoiseuilleInletProfile>(
      o::FaceDir::kYMin, 0.1, 0.9, 0.1, 0.9, 0.05, std::array<double, 3>{0.0, 1.0, 0.0}, 10.0);
  o::gpu::ConcreteDomainBoundaryHandler bc(blocks, faces.tree_boundary_faces(), specs, n, n, n, omega);
  o::DomainBoundaryHandler& base = bc;  // still usable wherever the reference's handler is
  base.set_time(0.0);
  o::gpu::TimeLoop loop(forest, blocks, ghosts, coupler, bc, omega, false,
                        o::ConstRhoStatsScope::kFluidAndBoundary);
  o::gpu::TimeLoop loop32(forest, blocks, ghosts, coupler, bc, omega, true, o::ConstRhoStatsScope::kFluidOnly, 32);
  o::NoOpDomainBoundaryHandler noop;
  o::gpu::TimeLoop plain(forest, blocks, ghosts, coupler, noop, omega);
  loop.set_snapshot_post_collide(true);
  for (int i = 0; i < 10; ++i) loop.advance_one();
  loop.advance(16);
  const o::TimeLoopCounters& c = loop.counters();
  double s = static_cast<double>(c.collide.size() + c.stream.size() + c.coupler_calls.size());
  s += loop.max_level() + (loop.snapshot_post_collide() ? 1 : 0) + loop.average_rho() + loop.mass();
  const std::array<double, 3> f = loop.drag(o::BoundingBox{0.9, 0.4, 0.0, 1.1, 0.6, 1.0});
  loop.sync_to_host();
  loop.sync_to_device();
  loop.reset_counters();
  (void)loop.device();
  return s + f[0] + loop32.average_rho() + plain.mass();
}
'''


def _flags():
    conda = Path("/opt/conda/include")
    inc = [f"-I{REF}", f"-I{REF / 'src' / 'solver' / 'lbm'}", f"-I{ROOT / 'include'}"]
    if (conda / "mpi.h").exists():
        inc.append(f"-I{conda}")
    return ["-std=c++20", "-fsyntax-only", "-Wall"] + inc


@pytest.mark.skipif(not (REF / "src" / "solver" / "lbm" / "time_loop" / "time_loop.h").exists(),
                    reason="reference sources not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_shim_compiles_against_reference_headers(tmp_path):
    src = tmp_path / "use_shim.cpp"
    src.write_text(USE)
    r = subprocess.run(["g++"] + _flags() + [str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "octlb_gpu_time_loop.h" not in r.stderr, r.stderr[-4000:]  # no warnings from the shim


@pytest.mark.skipif(not (REF / "src" / "solver" / "lbm" / "time_loop" / "time_loop.h").exists(),
                    reason="reference sources not present")
def test_shim_rejects_wrong_constructor_arguments(tmp_path):
    # the constructor is TimeLoop's: a FacePairList in the ghosts' place does not compile
    bad = USE.replace("o::gpu::TimeLoop plain(forest, blocks, ghosts,", "o::gpu::TimeLoop plain(forest, blocks, faces,")
    src = tmp_path / "bad.cpp"
    src.write_text(bad)
    r = subprocess.run(["g++"] + _flags() + [str(src)], capture_output=True, text=True)
    assert r.returncode != 0
