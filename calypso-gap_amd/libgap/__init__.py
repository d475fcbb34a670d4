"""libgap -- GAP energy, forces and stress on an AMD Instinct MI355X.

Drop-in for the ``libgap`` package of gappy (gappy/libgap/__init__.py):

  * ``libgap.libgap``  the f2py extension module (fgap_calc, fgap_read,
    fget_bond with the reference's signatures, plus fgap_calc_status,
    fgap_calc_batch, fgap_read_dims/fgap_read2, fgap_descriptors);
  * ``libgap.GAP.Calculator``  the Python calculator;
  * ``libgap.BOND.Bond``  the minimum-bond helper.

The numerics run in HIP kernels on the GPU (built in-tree by
calypso-gap_amd/build.py); there is no CPU fallback.
"""
__version__ = "0.0.1"
