"""CPU FP64 oracle for the frequency-domain likelihood hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``pespace_amd``) may import,
call or execute anything in this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker (or the timed CPU baseline), never as the thing
that is measured or shipped.

The oracle is a plain numpy restatement of nnrui/pespace's FD path:

* ``oracle.detector``   grid, PSD, noise, orbits, polarization, Marsat-Baker
  single links, FD Michelson TDI, Whittle lnL
  (pespace/detector/{tdi,noise,orbit,antenna}.py, pespace/inference/likelihood_ti.py)
* ``oracle.phenom``     IMRPhenomD / PhenomHM amplitude, phase and t(f), with
  node tables and cubic-Hermite node->grid interpolation.  The reference holds
  no PhenomD/HM code (it imports the absent tiwave / ripplegw packages), so the
  waveform is the build's own restatement of the published model; see the
  module header for which of its functions are pinned and how.
* ``oracle.qnm``        Kerr quasi-normal-mode frequencies (Leaver's continued
  fraction) used for the ringdown of every mode.
* ``oracle.likelihood`` end-to-end lnL / response for a parameter set, and the
  central finite-difference gradient used to check the adjoint.
"""
