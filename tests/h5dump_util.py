"""Reading HDF5 files back with the HDF5 project's h5dump tool (independent of the writer)."""
import os
import re
import shutil
import subprocess

import numpy as np

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def h5_attr(path, name):
    """Value of one attribute: a number, a string, or a list of numbers (compound)."""
    out = subprocess.run([H5DUMP, "-a", name, "-m", "%.17g", path], capture_output=True, text=True, check=True).stdout
    body = out[out.index("DATA {"):]
    m = re.search(r'\(0\): "(.*)"', body)
    if m:
        return m.group(1)
    nums = [float(x) for x in re.findall(r"[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", body.split("(0):", 1)[1])]
    return nums[0] if len(nums) == 1 else nums


def h5_data(path, name, dtype, tmp):
    """Raw values of one dataset (h5dump's binary output, little endian)."""
    out = os.path.join(str(tmp), "ds.bin")
    subprocess.run([H5DUMP, "-d", name, "-b", "LE", "-o", out, path], capture_output=True, check=True)
    return np.fromfile(out, dtype=dtype)


def h5_ints(path, name):
    """Integers of a dataset from h5dump's text output (compound datasets have no binary form)."""
    out = subprocess.run([H5DUMP, "-d", name, "-y", "-w", "0", path], capture_output=True, text=True,
                         check=True).stdout
    return [int(x) for x in re.findall(r"-?\d+", out[out.index("DATA {"):])]


def h5_objects(path):
    out = subprocess.run([H5DUMP, "-n", path], capture_output=True, text=True, check=True).stdout
    return set(re.findall(r"(?:group|dataset)\s+(\S+)", out))
