"""Host emulation of the multi-model batch (CPU only).

tests/native/batch_emul.cpp links rustb_amd/csrc/batch.cpp and workspace.cpp,
unchanged, against stand-ins for the HIP runtime and the device stages, built
with AddressSanitizer and UndefinedBehaviorSanitizer: planning into passes and
portions, the arena layout, the work-queue tile table (emulated with the
kernel's thread mapping: every element written exactly once), graph capture
and replay, the eager launch mode, pass splits down to one k-point pair, the
per-size variant table and the FP64 re-solve of flagged k-points on both
Fourier routes.  Each model's conductivity of k-dependent stand-ins must equal
the same stand-ins summed over its own mesh (1e-12 relative)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_batch_host_emulation_under_sanitizers(tmp_path):
    exe = str(tmp_path / "batch_emul")
    src = [os.path.join(ROOT, "rustb_amd", "csrc", "batch.cpp"), os.path.join(ROOT, "rustb_amd", "csrc", "workspace.cpp"),
           os.path.join(HERE, "native", "batch_emul.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", *src, "-o", exe, "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "all cases ok" in r.stdout


def test_model_batch_argument_checks():
    """ModelBatch rejects bad arguments before any device work."""
    import numpy as np
    import pytest

    import rustb_amd as ra

    m = ra.synth_model(3, 2, False, 4, 2, 1, n_R_target=13)
    with pytest.raises(ra.TbError):
        ra.ModelBatch([], [])
    with pytest.raises(ra.TbError):
        ra.ModelBatch([m], [(4, 4), (4, 4)])
    with pytest.raises(ra.TbError):
        ra.ModelBatch([m], [(4, 4, 4)])
    with pytest.raises(ra.TbError):
        ra.ModelBatch([m], [(4, 0)])
    with pytest.raises(ra.TbError):
        ra.ModelBatch([m], [(4, 4)], directions=[np.eye(3)])
