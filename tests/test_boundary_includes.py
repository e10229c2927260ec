"""The drop-in header compiles in a translation unit shaped like the reference's
main.cpp (src/main.cpp:9-12, 70-74): the reference's NaiveSimulation header and
nbody/hip_simulation.h together, with either include directory first.  With the
reference's include/ first, "nbody/simulation.h" resolves to the reference's
Simulation (simulation.h:6-37), which has no virtual destructor and does not
include <vector>.  Compile-only (g++ -fsyntax-only); needs the reference tree,
so it is skipped where that is absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/include"

MAIN = r"""
#include <iostream>
#include <vector>
#include "nbody/naive_simulation.h"
#include "nbody/hip_simulation.h"
using Simulation = nbody::HipSimulation;   // was nbody::OpenClSimulation (main.cpp:12)
int main() {
    Simulation::Vector bounds = {1.0f, 1.0f, 1.0f, 0.0f};
    std::vector<Simulation::Particle> particles;
    particles.push_back(Simulation::Particle(bounds, bounds, 1.0f, 1.0f));
    Simulation simulation(bounds, particles, 0.001f, std::cout);
    Simulation::Scalar t = simulation.step();
    std::vector<Simulation::Particle> out = simulation.particles();
    std::vector<nbody::NaiveSimulation::Particle> naive_particles;
    nbody::NaiveSimulation naive(naive_particles, 1.0f, 0.001f);
    (void)t; (void)out; (void)naive;
    return 0;
}
"""


@pytest.mark.skipif(not os.path.isdir(REF_INC) or shutil.which("g++") is None,
                    reason="needs the reference's include/ and g++")
@pytest.mark.parametrize("order", ["reference_first", "repo_first"])
def test_main_shaped_tu_compiles(tmp_path, order):
    src = tmp_path / "main.cpp"
    src.write_text(MAIN)
    inc = [REF_INC, os.path.join(ROOT, "include")]
    if order == "repo_first":
        inc.reverse()
    cmd = ["g++", "-std=c++14", "-fsyntax-only"] + [f"-I{d}" for d in inc] + [str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
