"""CPU checkers for the signed-distance path -- TEST INFRASTRUCTURE ONLY.

``oracle.port``  the C restatement (liboracle.so) of the reference algorithm.
``oracle.ref``   the reference library itself, compiled from its headers into
                 oracle/_ref/ (present only where the reference sources were available
                 at build time).

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py import this
package, and only as the checker. The product (ebgeometry_amd) never imports it.
"""
