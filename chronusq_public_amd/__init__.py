"""MI355X-native direct-SCF G[D] (ChronusQ ``AOIntegrals::newTwoEContractDirect``)."""
