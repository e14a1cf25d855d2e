"""binsim_amd -- MI355X-native on-the-fly FCI / VQE operator path (libhipotf.so + host mirror)."""
