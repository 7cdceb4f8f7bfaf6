"""Write the golden fixtures under tests/golden/ from the reference's own data.

Run once in the build container (the reference tree is not present on the GPU
box; the fixtures it writes are committed):

    python tests/golden/make_golden.py /root/reference

Fixtures (data only -- inputs and expected outputs):

* rankine_pair.npz        -- resources/fftBenchmark/rankine_vortex01_{0,1}.tif decoded
                             to uint8 (512 x 512, 8-bit PackBits TIFFs).
* rankine_vortex01_{0,1}.tif -- the two TIFF files themselves (input of the image
                             reader's test).
* gpu_fft_golden.npz      -- resources/fftBenchmark/gpuFFTBenchmarkValidationData.matFloat:
                             1024 correlation planes of 31 x 31 produced by the
                             reference's GPU job over 16 x 16 IAs without overlap
                             (GpuFFTBenchmarkJob.java:161-245, tolerance 1e-2 :329).
* gauss1d_cases.json      -- (uX, uY, sigmaX, sigmaY, DIM, tolerance) of every @Test in
                             tests/.../interpolators/Gaussian1DInterpolatorAdvanced*Test.java;
                             the plane generator is GaussianTestHelper.java:27-30.
* list_kats.json          -- operation sequences and expected lists of
                             tests/.../util/SimpleFixedLengthFloatLinkedListTests.java.
* multipeak_kat.json      -- the 10 x 10 plane and expectations of
                             tests/.../maximum/FindMaximumMultiPeakTests.java:25-160.
* fastreal_kat.json       -- the 4 x 4 tiles of tests/.../xcorr/FastRealFFTXCorrTests.java:31-51.
* gauss_filter_golden.npz -- testFiles/GaussianFilter2D_{3px_2.0f,5px_0.48f}_rankine_vortex01_*.matFloat
                             (GaussianFilter2DTest.java:31-85, tolerances 4e-5 / 1e-4): the
                             filtered rankine_vortex01_0/1 images, kept as the 16-pixel border
                             strips (where the padding rules act) and a central 64 x 64 block.

GPL-3.0 data from QuickLabPIV-ng (Luis Mendes); attribution kept here.
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_matfloat(path):
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    from oracle import read_matfloat as rm  # noqa: E402

    return rm(path)


def gauss_cases(ref):
    tdir = os.path.join(ref, "tests/pt/quickLabPIV/interpolators")
    cases = []
    for fn in sorted(os.listdir(tdir)):
        m = re.match(r"Gaussian1DInterpolatorAdvanced(\d)Pixels(\w+)Test\.java", fn)
        if not m:
            continue
        pixels, variant = int(m.group(1)), m.group(2)
        src = open(os.path.join(tdir, fn)).read()
        dim = int(re.search(r"int DIM\s*=\s*(\d+)", src).group(1))
        kind = "guo" if variant == "HongWeiGuo" else "gauss1d"
        for body in re.split(r"@Test", src)[1:]:
            name = re.search(r"public void (\w+)\(", body).group(1)

            def val(var):
                expr = re.search(r"float " + var + r"\s*=\s*([-0-9.f/*+() ]+);", body).group(1)
                expr = re.sub(r"([0-9.])f", r"\1", expr)
                assert re.fullmatch(r"[-0-9./*+() ]+", expr), expr
                return float(np.float32(eval(expr)))  # literal arithmetic only

            tol = re.findall(r"result\.getMainPeak[IJ]\(\),\s*([0-9.e-]+)f?\)", body)
            cases.append(dict(file=fn, test=name, kind=kind, pixels=pixels, dim=dim,
                              uX=val("uX"), uY=val("uY"), sigmaX=val("sigmaX"),
                              sigmaY=val("sigmaY"), tol=float(tol[0]) if tol else 8e-3))
    return cases


# SimpleFixedLengthFloatLinkedListTests.java: (maxSize, kernelWidth, ops, expected list in
# iteration (ascending) order).  ops: [0, v, i, j] = addInAscendOrder, [1,0,0,0] = remove,
# [2,0,0,0] = clear.  The three tests that use the plain append add() are not on the
# MultiPeaks path and are left out.
LIST_KATS = [
    ("simpleThreeInsertInAscendingOrder :100-130", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9]], [[30, 3, 1], [50, 8, 9], [70, 1, 5]]),
    ("simpleFourInsertInAscendingOrderWithDiscard1 :133-165", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 55, 8, 8]],
     [[30, 3, 1], [55, 8, 8], [70, 1, 5]]),
    ("simpleFourInsertInAscendingOrderWithDiscard2 :168-200", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 35, 4, 1]],
     [[35, 4, 1], [50, 8, 9], [70, 1, 5]]),
    ("simpleFourInsertInAscendingOrderWithDiscard3 :203-233", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 75, 1, 6]],
     [[30, 3, 1], [50, 8, 9], [75, 1, 6]]),
    ("simpleFourInsertInAscendingOrderWith2Discard1 :236-264", 3, 5,
     [[0, 70, 7, 1], [0, 30, 3, 3], [0, 50, 8, 9], [0, 55, 5, 6]],
     [[55, 5, 6], [70, 7, 1]]),
    ("simpleFourInsertInAscendingOrderSkip1 :267-299", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 49, 7, 9]],
     [[30, 3, 1], [50, 8, 9], [70, 1, 5]]),
    ("simpleFourInsertInAscendingOrderSkip2 :302-334", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 29, 3, 2]],
     [[30, 3, 1], [50, 8, 9], [70, 1, 5]]),
    ("simpleFourInsertInAscendingOrderSkip3 :337-369", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 69, 2, 4]],
     [[30, 3, 1], [50, 8, 9], [70, 1, 5]]),
    ("simpleFourInsertInAscendingOrderSkip4 :371-403", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [0, 70, 2, 4]],
     [[30, 3, 1], [50, 8, 9], [70, 1, 5]]),
    ("simpleItemInsertionAndRemovalTest :405-415", 3, 5,
     [[0, 70, 2, 3], [1, 0, 0, 0]], []),
    ("clearAllTest :417-433", 3, 5,
     [[0, 70, 1, 5], [0, 30, 3, 1], [0, 50, 8, 9], [2, 0, 0, 0]], []),
]


def multipeak_kat():
    m = np.full((10, 10), 12.0, np.float32)
    m[3, 6] = 15.0
    m[4, 5] = 16.0
    m[9, 4] = 13.0
    return dict(source="FindMaximumMultiPeakTests.java:25-160", numberOfPeaks=4, kernelWidth=3,
                plane=m.tolist(),
                expect=[[4, 5, 16.0], [9, 4, 13.0], [None, None, 12.0]], tol=1e-6)


def fastreal_kat():
    F = np.zeros((4, 4), np.float32)
    G = np.zeros((4, 4), np.float32)
    F[0, :3] = [6, 6, 6]
    F[1, :3] = [7, 6, 6]
    F[2, :3] = [7, 7, 6]
    G[0, :3] = [5, 5, 6]
    G[1, :3] = [6, 6, 7]
    G[2, :3] = [7, 7, 7]
    return dict(source="FastRealFFTXCorrTests.java:31-51,195-199", F=F.tolist(), G=G.tolist(),
                tol=1e-3)


def copy_tiffs(ref):
    for k in (0, 1):
        name = f"rankine_vortex01_{k}.tif"
        shutil.copyfile(os.path.join(ref, "resources/fftBenchmark", name), os.path.join(HERE, name))


def main(ref):
    from PIL import Image

    copy_tiffs(ref)
    fa = os.path.join(ref, "resources/fftBenchmark/rankine_vortex01_0.tif")
    fb = os.path.join(ref, "resources/fftBenchmark/rankine_vortex01_1.tif")
    ia = np.array(Image.open(fa))
    ib = np.array(Image.open(fb))
    assert ia.dtype == np.uint8 and ia.shape == (512, 512), (ia.dtype, ia.shape)
    np.savez_compressed(os.path.join(HERE, "rankine_pair.npz"), a=ia, b=ib)

    names, mats = read_matfloat(os.path.join(
        ref, "resources/fftBenchmark/gpuFFTBenchmarkValidationData.matFloat"))
    planes = np.stack(mats).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "gpu_fft_golden.npz"), planes=planes,
                        names=np.array(names))

    with open(os.path.join(HERE, "gauss1d_cases.json"), "w") as f:
        json.dump(gauss_cases(ref), f, indent=0)
    with open(os.path.join(HERE, "list_kats.json"), "w") as f:
        json.dump([dict(source="SimpleFixedLengthFloatLinkedListTests." + s, maxSize=a, kernelWidth=b,
                        ops=o, expect=e) for (s, a, b, o, e) in LIST_KATS], f, indent=0)
    with open(os.path.join(HERE, "multipeak_kat.json"), "w") as f:
        json.dump(multipeak_kat(), f)
    with open(os.path.join(HERE, "fastreal_kat.json"), "w") as f:
        json.dump(fastreal_kat(), f)
    gauss = {}
    for key, fn in (("g3_2p0", "GaussianFilter2D_3px_2.0f_rankine_vortex01_0.matFloat"),
                    ("g5_0p48", "GaussianFilter2D_5px_0.48f_rankine_vortex01_0_1.matFloat")):
        _, mats = read_matfloat(os.path.join(ref, "testFiles", fn))
        for k, m in enumerate(mats):
            assert m.shape == (512, 512)
            gauss[f"{key}_{k}_top"] = m[:16]
            gauss[f"{key}_{k}_bottom"] = m[-16:]
            gauss[f"{key}_{k}_left"] = m[:, :16]
            gauss[f"{key}_{k}_right"] = m[:, -16:]
            gauss[f"{key}_{k}_centre"] = m[224:288, 224:288]
    np.savez_compressed(os.path.join(HERE, "gauss_filter_golden.npz"), **gauss)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
