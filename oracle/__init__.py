"""CPU oracle (test infrastructure only; never imported by the qsft_amd package)."""
