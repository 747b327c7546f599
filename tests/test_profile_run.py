"""The counter summaries of walicxe3d_amd/profile_run.py: the memory
counters are scaled by the 1 GiB device copy that ends every profiled run.
The state upload is a copy dispatch of the same grid size, so the copy is
told apart by dispatch order (the last of the largest grid)."""
import csv
import os

from walicxe3d_amd import profile_run as PR


def _write(path, rows):
    cols = ["Dispatch_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(cols, r)))


def test_calibration_copy_is_the_last_large_copy(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    # upload of a 4.2 GiB state (dispatch 3), small copies, the stage
    # kernels, then the 1 GiB calibration copy (dispatch 22), same grid as
    # the upload
    _write(str(d / "run_counter_collection.csv"), [
        (1, 512, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 0.78125),
        (3, 131072, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 4409600.0),
        (4, 512, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 0.0625),
        (9, 7000000, "void wx::k_stage<5, true, true, float, float, 4, -100>(wx::FusedArgs)", "WRITE_SIZE", 1.0e7),
        (22, 131072, "__amd_rocclr_copyBuffer", "WRITE_SIZE", 1048576.0),
    ])
    acc = PR._per_kernel(PR._read(str(d)), "WRITE_SIZE")
    assert PR._copy_reading(acc) == 1048576.0
    # the scale the summary applies: bytes per counter byte
    assert PR.COPY_BYTES / (1024.0 * PR._copy_reading(acc)) == 1.0
