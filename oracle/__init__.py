"""The oracle: a CPU restatement of pywave's acoustic FD engine and FWI gradient.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import this package, and only as the checker (never as the thing
measured or shipped).  The product package ``pywave_amd`` never imports it.

Contents
  orc_afd.c / orc_fwi.c  plain-C restatement (built into oracle/liborc.so by oracle/Makefile)
  oracle.py              ctypes bindings + numpy pieces (wavelet, geometry)
  refrun.py              runner for oracle/_ref/pwref, the reference's own C compiled from
                         its sources (golden vectors, CPU baseline)

Parity status of each oracle function (what pins it) is listed in oracle/oracle.py.
"""
