"""The drop-in's headers keep the reference's struct layouts (callers read the
fields directly, so the layout is API), and the ctypes mirror used by the GPU
tests agrees with what the C compiler lays out.  CPU only."""
import ctypes as C
import os
import subprocess

import pytest

from dcell_amd import dropin as DI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "LevelSetMethod.h"
#include "LSM_private.h"
#include "ImmersedInterfaceMethod.h"
#define F(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(struct T, f))
int main(void) {
  printf("Grid %zu\nLevelSet %zu\nIrregularNode %zu\nParticle %zu\nParticleLS %zu\nIIM %zu\n",
         sizeof(struct _Grid), sizeof(struct _LevelSet), sizeof(struct _IrregularNode),
         sizeof(struct _Particle), sizeof(struct _ParticleLS), sizeof(struct _IIM));
  F(_Grid, p); F(_Grid, n); F(_Grid, d); F(_Grid, is2D); F(_Grid, v1); F(_Grid, dof);
  F(_Grid, SIZE); F(_Grid, Interpolate);
  F(_LevelSet, band); F(_LevelSet, CFLcount); F(_LevelSet, AdvectCount); F(_LevelSet, maxVel);
  F(_LevelSet, pls); F(_LevelSet, Advect);
  F(_IrregularNode, nx); F(_IrregularNode, k); F(_IrregularNode, X); F(_IrregularNode, uj);
  F(_ParticleLS, S_INIT); F(_ParticleLS, AdvectParticles); F(_ParticleLS, particles0);
  F(_IIM, Np); F(_IIM, eps); F(_IIM, F); F(_IIM, context); F(_IIM, dh); F(_IIM, sidx);
  return 0;
}
"""


def test_struct_layouts(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT}/include/dcell", str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.split("\n") if line)
    out = {k: int(v) for k, v in out.items()}
    # sizes fixed by the reference's field lists on LP64 with 32-bit PetscInt
    assert out["Grid"] == 4096 + 8 + 24 + 24 + 8 + 8 + 8 + 12 + 4 + 16 + 8
    assert out["Particle"] == 32
    assert out["IrregularNode"] == C.sizeof(DI.IrregularNodeS)
    mirror = {"Grid": DI.GridS, "LevelSet": DI.LevelSetS, "ParticleLS": DI.ParticleLSS, "IIM": DI.IIMS}
    for name, cls in mirror.items():
        assert out[name] == C.sizeof(cls), name
    for key, off in out.items():
        if "." not in key:
            continue
        struct, field = key.split(".")
        cls = {"_Grid": DI.GridS, "_LevelSet": DI.LevelSetS, "_IrregularNode": DI.IrregularNodeS,
               "_ParticleLS": DI.ParticleLSS, "_IIM": DI.IIMS}[struct]
        assert getattr(cls, field).offset == off, key


def test_dropin_libraries_export_reference_names():
    if not os.path.exists(DI.LIB_LEVELSET):
        pytest.fail("libdcell_levelset.so not built")
    syms = subprocess.run(["nm", "-D", "--defined-only", DI.LIB_LEVELSET], capture_output=True, text=True,
                          check=True).stdout
    for name in ("LevelSetAdvectSL", "LevelSetAdvectSL_2D", "LevelSetAdvectSL_3D",
                 "LevelSetAdvectSLRK2HalfStep", "LevelSetAdvectSLRK2FullStep", "LevelSetAdvectAndReinit",
                 "LevelSetAdvectPLS", "LevelSetAdvectPLSRK2HalfStep", "LevelSetAdvectPLSRK2FullStep",
                 "ParticleLS_AdvectParticles2D", "ParticleLS_AdvectParticles3D",
                 "ParticleLS_ErrorCorrection2D", "ParticleLS_ErrorCorrection3D", "ParticleLS_AdjustRadii",
                 "LevelSetCFLIncrement", "LevelSet_MaxVelocity", "LevelSetCurvature",
                 "LevelSetUpdateIrregularNodeList", "LevelSetUpdateIrregularNodeList_2D", "LevelSetReinitialize",
                 "LevelSetUpdateIrregularNodeList_3D", "IIMUpdateSurfaceQuantities_2D",
                 "IIMUpdateSurfaceQuantities_3D", "LevelSetAdvectSLSetScheme", "LevelSetPLSSetReferenceCornerZ", "LevelSetGPUFinalize",
                 "LevelSetGPUTransferStats"):
        assert f" T {name}\n" in syms, name
