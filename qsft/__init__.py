"""The reference's import paths (qsft.qsft, qsft.query, ...) served by qsft_amd.

A script written for basics-lab/qsft (for example one that does
`from qsft.qsft import QSFT`) runs unchanged with this repository's root on
sys.path: every module here re-exports the qsft_amd module of the same role.
"""
