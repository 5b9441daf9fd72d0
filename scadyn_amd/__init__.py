"""scadyn_amd: MI355X-native JVIE T-matrix solve (precorrected-FFT GMRES)."""
