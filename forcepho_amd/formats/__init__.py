"""Readers of the data formats forcepho's patches are built from: FITS images
and binary tables (``fits``), HDF5 PSF stores and spline tables (``hdf5``),
and the TAN world coordinate systems of image headers (``wcs``)."""
