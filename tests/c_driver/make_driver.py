"""Compile the link-level drop-in driver (dropin_driver.c) with gcc against
include/orth.h and liborth_hip.so (rpath to the library's directory)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIBDIR = os.path.join(ROOT, "one-synchronization-bcgs_amd", "lib")
OUT = os.path.join(HERE, "_build", "dropin_driver")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-Wall", "-std=c99", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "dropin_driver.c"), "-L" + LIBDIR, "-lorth_hip",
           # The library's directory, relative to the driver's.
           "-Wl,-rpath,$ORIGIN/" + os.path.relpath(LIBDIR, os.path.dirname(OUT)), "-lm",
           "-o", OUT]
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build())
