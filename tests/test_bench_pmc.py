"""bench.py's counter summary: the raster/shade traffic from rocprofv3 --pmc
output, calibrated by the streaming kernel of known size (host logic only)."""
import csv
import os

import bench

HEADER = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def write_run(d, counter, rows):
    os.makedirs(os.path.join(d, "host"), exist_ok=True)
    with open(os.path.join(d, "host", "123_counter_collection.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(HEADER)
        for i, (name, value) in enumerate(rows):
            w.writerow([i, name, counter, value])


def test_pmc_summary_calibrates_and_sums(tmp_path):
    raster = "void edxr::k_raster<false>(edxr::BatchParams, edxr::FrameParams const*, edxr::DeviceBuffers)"
    shade = "edxr::k_shade(edxr::BatchParams, edxr::FrameParams const*, edxr::DeviceBuffers)"
    neg = "void at::native::vectorized_elementwise_kernel<4, at::native::neg_kernel_cuda(...)>"
    other = "edxr::k_bin_scatter(edxr::BatchParams, edxr::FrameParams const*, edxr::DeviceBuffers)"
    known = bench.CALIBRATION_BYTES / 1024
    # the fetch counter reads half the bytes, the write counter all of them
    write_run(str(tmp_path / "f"), "FETCH_SIZE",
              [(raster, 100.0), (raster, 300.0), (shade, 50.0), (neg, known / 2), (other, 9999.0)])
    write_run(str(tmp_path / "w"), "WRITE_SIZE", [(raster, 40.0), (shade, 10.0), (neg, known), (other, 1.0)])
    s = bench.pmc_summary(str(tmp_path / "f"), str(tmp_path / "w"))
    assert abs(s["fetch_factor"] - 2.0) < 1e-12 and abs(s["write_factor"] - 1.0) < 1e-12
    assert s["fetch_kib"]["k_raster<false>"] == 200.0
    assert s["dispatches"]["fetch"]["k_raster<false>"] == 2
    # (200 + 50) KiB x 2 + (40 + 10) KiB x 1
    assert s["traffic_bytes"] == 550 * 1024
    assert sorted(s["kernels"]) == ["k_raster<false>", "k_shade"]
