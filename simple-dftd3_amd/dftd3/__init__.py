"""MI355X implementation of the DFT-D3 dispersion energy, gradient and virial.

``dftd3.interface`` is the drop-in replacement of simple-dftd3's Python
interface for the energy / gradient / virial path; ``dftd3.torch_api`` adds
device-tensor entry points (batches, autograd).
"""

__version__ = "1.6.0"
