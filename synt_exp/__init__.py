"""The reference's synt_exp package path, served by qsft_amd."""
