"""Random streams (randomGen.pyx:22-35 replaced by a counter-based generator).

The reference draws everything from one global GSL MT19937 stream, so its results depend on
the order of calls.  Here every draw is Philox4x32-10 (Salmon et al., SC'11) addressed by a
(seed, item, slot, step, purpose) counter, so device threads draw independently and results
do not depend on launch order.  seed(x) resets the key exactly like randomGen.seed.
"""
import time

_state = {'seed': int(time.time() * 256) & 0xFFFFFFFFFFFFFFFF, 'calls': 0}


def seed(x):
    """randomGen.pyx:25-26."""
    _state['seed'] = int(x) & 0xFFFFFFFFFFFFFFFF
    _state['calls'] = 0


def key():
    """Philox key (k0, k1) of the current seed."""
    s = _state['seed']
    return s & 0xFFFFFFFF, (s >> 32) & 0xFFFFFFFF


def next_stream():
    """A fresh stream index for draws that are not tied to a time step."""
    _state['calls'] += 1
    return _state['calls']
