"""Python mirror of the reference's museqa.sequence / museqa.database /
museqa.pairwise modules (src/python/*.pyx) plus museqa.phylogeny, over the
C++ host layer (msa_amd/csrc/host) and the HIP kernels behind libmsa_hip.so."""
from . import database, pairwise, phylogeny, sequence  # noqa: F401
